// VTU/PVTU output in the reference's file format (io/Output.h:25-175,
// io/VTK.h:316-560, 854-1355): one voxel (3D) or pixel (2D) cell per
// position, corner points shared between cells and numbered in order of
// first use, binary appended data.  A variable is a host array that the
// writer reads at every write (the reference's Output keeps references to
// its fields the same way).
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/badchimp_amd.h"
#include "errors.hpp"

using lbamd::fail_global;

namespace {

struct Var {
  std::string name;
  int dim = 1;           // components per cell in the data (1 or nd)
  const void* data = nullptr;
  int comps() const { return dim == 2 ? 3 : dim; }  // 2-vectors are written with a zero third component
};

struct File {
  std::string name;
  std::vector<Var> vars;
  int nwrite = 0;
};

size_t esize(int dtype) { return dtype == LB_VTU_F32 || dtype == LB_VTU_I32 ? 4 : 8; }
const char* tname(int dtype) {
  return dtype == LB_VTU_F64 ? "Float64" : (dtype == LB_VTU_F32 ? "Float32" : "Int32");
}

bool make_dir(const std::string& d) {
  struct stat st;
  if (stat(d.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
  return mkdir(d.c_str(), 0700) == 0 || errno == EEXIST;
}

// Name="x" type="T" NumberOfComponents="n" format="appended" offset="o"
std::string dataarray(const std::string& name, const char* type, int comps, int64_t offset) {
  return "<DataArray Name=\"" + name + "\" type=\"" + type + "\" NumberOfComponents=\"" + std::to_string(comps) +
         "\" format=\"appended\" offset=\"" + std::to_string(offset) + "\" > </DataArray>";
}

}  // namespace

struct lb_vtu {
  int nd = 3, dtype = LB_VTU_F64;
  int rank = 0, nprocs = 1;
  int32_t n_cells = 0;
  std::string dir;  // ends with '/'
  // the mesh (VTK.h:395-555)
  std::vector<float> points;   // [n_points*3]
  std::vector<int32_t> conn;   // [n_cells*corners]
  int corners = 8, cell_type = 11;
  int64_t grid_bytes = 0;      // appended bytes of points and cells, with their length words
  std::vector<File> files;
};

namespace {

// Mesh::set_size / get_stride / calc_points_and_conn (VTK.h:460-555).
void build_mesh(lb_vtu& v, const int32_t* pos) {
  const int nd = v.nd;
  static const int vox[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  v.corners = nd == 3 ? 8 : 4;
  v.cell_type = nd == 3 ? 11 : 8;
  std::vector<int64_t> mn(nd, 999999), mx(nd, -1);
  for (int64_t k = 0; k < static_cast<int64_t>(v.n_cells) * nd; ++k) {
    const int i = static_cast<int>(k % nd);
    mx[i] = std::max<int64_t>(mx[i], pos[k]);
    mn[i] = std::min<int64_t>(mn[i], pos[k]);
  }
  std::vector<int64_t> stride(nd, 1);
  for (int i = 1; i < nd; ++i) stride[i] = stride[i - 1] * ((mx[i - 1] - mn[i - 1] + 1) + 1);
  const size_t nc = static_cast<size_t>(v.n_cells) * v.corners;
  std::vector<int64_t> index(nc);
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  for (int32_t n = 0; n < v.n_cells; ++n)
    for (int p = 0; p < v.corners; ++p) {
      int64_t idx = 0;
      for (int i = 0; i < nd; ++i) idx += (pos[static_cast<int64_t>(n) * nd + i] + vox[p][i]) * stride[i];
      index[static_cast<size_t>(n) * v.corners + p] = idx;
      lo = std::min(lo, idx);
      hi = std::max(hi, idx);
    }
  v.conn.assign(nc, 0);
  v.points.clear();
  if (nc == 0) return;
  std::vector<int32_t> first(static_cast<size_t>(hi - lo + 1), -1);
  int32_t n_pts = 0;
  for (size_t k = 0; k < nc; ++k) {
    int32_t& slot = first[static_cast<size_t>(index[k] - lo)];
    if (slot < 0) {
      slot = n_pts++;
      const int64_t n = static_cast<int64_t>(k / v.corners);
      const int p = static_cast<int>(k % v.corners);
      // corner coordinates shifted by half a cell so the node sits at the cell centre
      for (int i = 0; i < 3; ++i)
        v.points.push_back(i < nd ? static_cast<float>(static_cast<double>(pos[n * nd + i] + vox[p][i]) - 0.5)
                                  : 0.0f);
    }
    v.conn[k] = slot;
  }
}

bool put(FILE* f, const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
bool put_block(FILE* f, const void* p, uint32_t nbytes) { return put(f, &nbytes, 4) && put(f, p, nbytes); }

int write_vtu(lb_vtu& v, File& fl) {
  char fname[4096];
  std::snprintf(fname, sizeof(fname), "%svtu/%04d_%s_%07d.vtu", v.dir.c_str(), v.rank, fl.name.c_str(), fl.nwrite);
  std::unique_ptr<FILE, int (*)(FILE*)> fh(std::fopen(fname, "wb"), std::fclose);
  if (!fh) return fail_global(LB_ERR_IO, "could not open %s for writing", fname);
  FILE* f = fh.get();
  const int64_t n_pts = static_cast<int64_t>(v.points.size() / 3);
  std::string h;
  h += "<?xml version=\"1.0\"?>\n";
  h += "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
  h += "  <UnstructuredGrid>\n";
  h += "    <Piece NumberOfPoints=\"" + std::to_string(n_pts) + "\" NumberOfCells=\"" + std::to_string(v.n_cells) +
       "\">\n";
  h += "      <Points>\n";
  int64_t off = 0;
  h += "        " + dataarray("points", "Float32", 3, off) + "\n";
  off += 4 + 4 * 3 * n_pts;
  h += "      </Points>\n      <Cells>\n";
  h += "        " + dataarray("connectivity", "Int32", 1, off) + "\n";
  off += 4 + 4 * static_cast<int64_t>(v.conn.size());
  h += "        " + dataarray("offsets", "Int32", 1, off) + "\n";
  off += 4 + 4 * static_cast<int64_t>(v.n_cells);
  h += "        " + dataarray("types", "Int32", 1, off) + "\n";
  off += 4 + 4 * static_cast<int64_t>(v.n_cells);
  h += "      </Cells>\n";
  std::string scal, vec;
  for (auto& x : fl.vars) {
    std::string& s = x.dim > 1 ? vec : scal;
    s += (s.empty() ? "" : ", ") + x.name;
  }
  h += "      <CellData Scalars=\"" + scal + "\" Vectors=\"" + vec + "\">\n";
  const size_t es = esize(v.dtype);
  for (auto& x : fl.vars) {
    h += "        " + dataarray(x.name, tname(v.dtype), x.comps(), off) + "\n";
    off += 4 + static_cast<int64_t>(es) * x.comps() * v.n_cells;
  }
  h += "      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n_";
  bool ok = put(f, h.data(), h.size());
  ok = ok && put_block(f, v.points.data(), static_cast<uint32_t>(4 * v.points.size()));
  ok = ok && put_block(f, v.conn.data(), static_cast<uint32_t>(4 * v.conn.size()));
  std::vector<int32_t> cells(v.n_cells);
  for (int32_t c = 0; c < v.n_cells; ++c) cells[c] = (c + 1) * v.corners;
  ok = ok && put_block(f, cells.data(), static_cast<uint32_t>(4 * cells.size()));
  std::fill(cells.begin(), cells.end(), v.cell_type);
  ok = ok && put_block(f, cells.data(), static_cast<uint32_t>(4 * cells.size()));
  std::vector<char> buf;
  for (auto& x : fl.vars) {
    const int comps = x.comps();
    buf.assign(es * comps * static_cast<size_t>(v.n_cells), 0);
    const char* src = static_cast<const char*>(x.data);
    for (int32_t c = 0; c < v.n_cells; ++c)
      std::memcpy(&buf[es * comps * static_cast<size_t>(c)], src + es * x.dim * static_cast<size_t>(c), es * x.dim);
    ok = ok && put_block(f, buf.data(), static_cast<uint32_t>(buf.size()));
  }
  ok = ok && put(f, "  </AppendedData>\n</VTKFile>\n", 29);
  ok = ok && std::fflush(f) == 0;
  if (!ok) return fail_global(LB_ERR_IO, "write error on %s", fname);
  return LB_OK;
}

// PVTU_file::write (VTK.h:1226-1355): rank 0 writes the header, every rank
// its 100-byte piece line at header + 100 rank, the last rank the footer.
int write_pvtu(lb_vtu& v, File& fl, double time) {
  char fname[4096], tbuf[64], piece[128], ptxt[4096];
  std::snprintf(fname, sizeof(fname), "%s%s_%07d.pvtu", v.dir.c_str(), fl.name.c_str(), fl.nwrite);
  const std::time_t now = std::time(nullptr);
  std::tm tm{};
  localtime_r(&now, &tm);
  std::snprintf(tbuf, sizeof(tbuf), "%02d.%02d.%02d %02d:%02d:%02d", tm.tm_mday, tm.tm_mon + 1, tm.tm_year + 1900,
                tm.tm_hour, tm.tm_min, tm.tm_sec);
  char tstr[64];
  std::snprintf(tstr, sizeof(tstr), "%g", time);
  std::string h;
  h += "<?xml version=\"1.0\"?>\n";
  h += std::string("<!-- Created ") + tbuf + " -->\n";
  h += std::string("<!-- time = ") + tstr + " s -->\n";
  h += "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
  h += "  <PUnstructuredGrid GhostLevel=\"0\">\n";
  h += "    <PPoints>\n      <PDataArray type=\"Float32\" NumberOfComponents=\"3\" />\n    </PPoints>\n";
  h += "    <PCellData>\n";
  h += "      <PDataArray type=\"Int32\" Name=\"connectivity\" />\n";
  h += "      <PDataArray type=\"Int32\" Name=\"offsets\" />\n";
  h += "      <PDataArray type=\"Int32\" Name=\"types\" />\n";
  h += "    </PCellData>\n    <PCellData>\n";
  for (auto& x : fl.vars)
    h += "      <PDataArray Name=\"" + x.name + "\" type=\"" + tname(v.dtype) + "\" NumberOfComponents=\"" +
         std::to_string(x.comps()) + "\" />\n";
  h += "    </PCellData>\n";
  std::snprintf(ptxt, sizeof(ptxt), "    <Piece Source=\"vtu/%04d_%s_%07d.vtu\" />", v.rank, fl.name.c_str(),
                fl.nwrite);
  std::snprintf(piece, sizeof(piece), "%-99.99s\n", ptxt);
  FILE* f = std::fopen(fname, v.rank == 0 ? "wb" : "r+b");
  if (!f) return fail_global(LB_ERR_IO, "could not open %s", fname);
  bool ok = true;
  if (v.rank == 0) ok = put(f, h.data(), h.size());
  ok = ok && std::fseek(f, static_cast<long>(h.size() + 100 * static_cast<size_t>(v.rank)), SEEK_SET) == 0;
  ok = ok && put(f, piece, 100);
  if (v.rank == v.nprocs - 1) ok = ok && put(f, "  </PUnstructuredGrid>\n</VTKFile>\n", 34);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? LB_OK : fail_global(LB_ERR_IO, "write error on %s", fname);
}

}  // namespace

extern "C" {

int lb_vtu_create(int32_t nd, const int32_t* pos, int32_t n_cells, const char* dir, int32_t rank, int32_t nprocs,
                  int32_t dtype, lb_vtu** out) {
  if (!out || !dir || (n_cells > 0 && !pos) || n_cells < 0 || (nd != 2 && nd != 3) || nprocs < 1 || rank < 0 ||
      rank >= nprocs || (dtype != LB_VTU_F64 && dtype != LB_VTU_F32 && dtype != LB_VTU_I32))
    return fail_global(LB_ERR_ARG, "lb_vtu_create: bad argument");
  try {
    auto v = std::make_unique<lb_vtu>();
    v->nd = nd;
    v->dtype = dtype;
    v->rank = rank;
    v->nprocs = nprocs;
    v->n_cells = n_cells;
    v->dir = dir;
    if (v->dir.empty() || v->dir.back() != '/') v->dir += '/';
    if (!make_dir(v->dir) || !make_dir(v->dir + "vtu/"))
      return fail_global(LB_ERR_IO, "could not create %svtu/", v->dir.c_str());
    build_mesh(*v, pos);
    *out = v.release();
  } catch (const std::bad_alloc&) {
    return fail_global(LB_ERR_NOMEM, "lb_vtu_create: out of host memory");
  }
  return LB_OK;
}

int lb_vtu_add_file(lb_vtu* v, const char* name) {
  if (!v || !name || !*name) return fail_global(LB_ERR_ARG, "lb_vtu_add_file: bad argument");
  v->files.push_back(File{name, {}, 0});
  return LB_OK;
}

int lb_vtu_add_variable(lb_vtu* v, const char* name, int32_t dim, const void* data) {
  if (!v || !name || !*name || (v->n_cells > 0 && !data) || (dim != 1 && dim != v->nd))
    return fail_global(LB_ERR_ARG, "lb_vtu_add_variable: bad argument (dim must be 1 or %d)", v ? v->nd : 3);
  if (v->files.empty())
    return fail_global(LB_ERR_STATE, "lb_vtu_add_variable: add an output file with lb_vtu_add_file first");
  v->files.back().vars.push_back(Var{name, dim, data});
  return LB_OK;
}

int lb_vtu_write(lb_vtu* v, double time) {
  if (!v) return fail_global(LB_ERR_ARG, "null argument");
  for (auto& fl : v->files) {
    int rc = write_vtu(*v, fl);
    if (rc == LB_OK) rc = write_pvtu(*v, fl, time);
    if (rc != LB_OK) return rc;
    ++fl.nwrite;
  }
  return LB_OK;
}

int32_t lb_vtu_num_points(const lb_vtu* v) { return v ? static_cast<int32_t>(v->points.size() / 3) : -1; }

int lb_vtu_destroy(lb_vtu* v) {
  delete v;
  return LB_OK;
}

}  // extern "C"
