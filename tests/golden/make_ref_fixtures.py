"""Run the compiled reference on small geometries and store its fields as fixtures.

Needs the reference binaries built by oracle/build_ref.sh (in the build
container only; /root/reference is not available elsewhere).  For every case:

  1. the geometry is written as .vtklb by the oracle's restatement of the
     reference writer (itself pinned byte-for-byte against files written by
     PythonScripts/vtklb.py, see make_vtklb_fixtures.py),
  2. an input.dat with the case parameters is written,
  3. the reference case (std_case D2Q9 / D3Q19, its two-relaxation-time
     variant, the generalized-Newtonian polymer_viscosity_tabulation,
     std_one_phase with tagged boundaries, or twophase) runs with
     iterations.write = iterations.max, so it writes VTK at i = 0 and i = max,
  4. the VTU of i = max is parsed: rho (rho0, rho1) and vel on the bulk nodes,
     in bulk order, i.e. the fields of the pre-collision state after max
     steps (std_case/main.cpp:109-155, main_TWOPHASE.cpp:236-424),
  5. inputs (geometry, parameters) and outputs are stored in ref_<case>.npz;
     for a few small cases the VTU and PVTU files themselves are kept in
     tests/golden/vtu/ as fixtures of the VTU writer.

Run:  python tests/golden/make_ref_fixtures.py [case ...]
"""
import glob
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "badchimp-cpp_amd"))

from oracle.vtklb_oracle import Geometry  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
VT = os.path.join(HERE, "vtklb")


def parse_vtu(path):
    """Cell data arrays of a VTU file written by the reference (VTK.h:1094-1180):
    raw appended data, each array preceded by its byte count (uint32)."""
    raw = open(path, "rb").read()
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">')
    base = tail.index(b"_") + 1
    out = {}
    for m in re.finditer(rb'<DataArray Name="([^"]+)" type="(\w+)" NumberOfComponents="(\d+)" '
                         rb'format="appended" offset="(\d+)"', head):
        name, typ, ncomp, off = m.group(1).decode(), m.group(2).decode(), int(m.group(3)), int(m.group(4))
        nbytes = int(np.frombuffer(tail, dtype=np.uint32, count=1, offset=base + off)[0])
        dt = {"Float64": np.float64, "Float32": np.float32, "Int32": np.int32}[typ]
        arr = np.frombuffer(tail, dtype=dt, count=nbytes // np.dtype(dt).itemsize,
                            offset=base + off + 4).copy()
        out[name] = arr.reshape(-1, ncomp) if ncomp > 1 else arr
    return out


def last_vtu(outdir, base):
    files = sorted(glob.glob(os.path.join(outdir, "vtu", "0000_%s_*.vtu" % base)))
    return files[-1]


def input_dat(max_it, fluid_lines, extra=""):
    return ("<iterations>\n    max %d\n    write %d\n<end>\n<fluid>\n%s<end>\n%s"
            % (max_it, max_it, "".join("    %s\n" % l for l in fluid_lines), extra))


def fmt(v):
    return " ".join(repr(float(x)) for x in np.atleast_1d(v))


def run_one_phase(geo, attrs, dat):
    """std_one_phase (oracle/_ref/one_phase): output under SPE2024/<directory>/,
    the mass fluxes in <directory>.flux in the working directory."""
    tmp = tempfile.mkdtemp(prefix="lbref_")
    try:
        g = Geometry(geo, "D3Q19", "xy")
        os.makedirs(os.path.join(tmp, "input", "mpi"))
        os.makedirs(os.path.join(tmp, "SPE2024"))
        os.makedirs(os.path.join(tmp, "run"))
        with open(os.path.join(tmp, "input", "mpi", "tmp0.vtklb"), "w") as fh:
            fh.write(g.vtklb_text(1, attrs))
        with open(os.path.join(tmp, "input", "input.dat"), "w") as fh:
            fh.write(dat)
        env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(REF, "mpilib"), OMP_NUM_THREADS="1")
        res = subprocess.run([os.path.join(REF, "one_phase")], cwd=os.path.join(tmp, "run"), env=env,
                             capture_output=True, text=True, timeout=3600)
        if res.returncode != 0:
            raise RuntimeError("one_phase failed: %s %s" % (res.stdout[-2000:], res.stderr[-2000:]))
        data = parse_vtu(last_vtu(os.path.join(tmp, "SPE2024", "run"), "lb_run"))
        text = open(os.path.join(tmp, "run", "run.flux")).read()
        q = re.findall(r"q(\d) = (\S+) \(", text)
        data["flux"] = np.array([float(q[-2][1]), float(q[-1][1])])
        return data
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def keep_output(outdir, base, name):
    """Copy the reference's last VTU file and its PVTU into tests/golden/vtu/
    (the byte-level fixtures of the VTU writer)."""
    dst = os.path.join(HERE, "vtu")
    os.makedirs(dst, exist_ok=True)
    vtu = last_vtu(outdir, base)
    shutil.copyfile(vtu, os.path.join(dst, name + ".vtu"))
    n = os.path.basename(vtu).split("_")[-1]  # NNNNNNN.vtu
    shutil.copyfile(os.path.join(outdir, "%s_%s.pvtu" % (base, n[:-4])), os.path.join(dst, name + ".pvtu"))


def run_case(binary, geo, lattice, periodic, attrs, dat, twophase=False, outsub="tp/", keep=None,
             extra_inputs=None):
    tmp = tempfile.mkdtemp(prefix="lbref_")
    try:
        g = Geometry(geo, lattice, periodic)
        if twophase:
            root = os.path.join(tmp, "run")
        else:
            root = tmp
        os.makedirs(os.path.join(root, "input", "mpi"))
        os.makedirs(os.path.join(root, "output", outsub) if twophase else os.path.join(root, "output"))
        with open(os.path.join(root, "input", "mpi", "tmp0.vtklb"), "w") as fh:
            fh.write(g.vtklb_text(1, attrs))
        with open(os.path.join(root, "input", "input.dat"), "w") as fh:
            fh.write(dat)
        for fname, text in (extra_inputs or {}).items():
            with open(os.path.join(root, "input", fname), "w") as fh:
                fh.write(text)
        cwd = root if twophase else os.path.join(tmp, "run")
        os.makedirs(cwd, exist_ok=True)
        env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(REF, "mpilib"), OMP_NUM_THREADS="1")
        res = subprocess.run([os.path.join(REF, binary)], cwd=cwd, env=env, capture_output=True,
                             text=True, timeout=3600)
        if res.returncode != 0:
            raise RuntimeError("%s failed: %s %s" % (binary, res.stdout[-2000:], res.stderr[-2000:]))
        if twophase:
            od = os.path.join(root, "output", outsub)
            data = parse_vtu(last_vtu(od, "fluid"))
            force = np.loadtxt(os.path.join(od, "force.dat"), ndmin=2)
            data["force_x"] = force
            if keep:
                keep_output(od, "fluid", keep)
        else:
            data = parse_vtu(last_vtu(os.path.join(tmp, "output"), "lb_run"))
            if keep:
                keep_output(os.path.join(tmp, "output"), "lb_run", keep)
        return data
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def load_geo(name):
    return np.load(os.path.join(VT, name + "_geo.npy"))


def sphere_geo(shape, radius, porosity, seed, periodic, walls):
    import lbamd
    geo, info = lbamd.sphere_pack(shape, radius, porosity, seed, periodic=periodic, walls=walls)
    return geo.astype(np.int64), info


def std_cases():
    """(name, lattice, binary, geo, periodic, attrs, tau, force, max_it, profile_only)"""
    g = load_geo("d2q9_channel_8x6_px")
    yield "d2q9_channel_8x6", "D2Q9", g, "x", [("init_rho", np.ones(g.shape))], 1.0, [1e-6, 0.0], 20, False
    g = load_geo("d2q9_random_7x5_pxy")
    r = np.load(os.path.join(VT, "d2q9_random_7x5_pxy_attr_init_rho.npy"))
    yield "d2q9_random_7x5", "D2Q9", g, "xy", [("init_rho", r)], 0.8, [1e-5, -3e-6], 50, False
    g = np.ones((64, 32), dtype=np.int64)
    g[:, 0] = 0
    g[:, -1] = 0
    yield "d2q9_channel_64x32", "D2Q9", g, "x", [("init_rho", np.ones(g.shape))], 0.7, [1e-5, 0.0], 500, False
    g = np.ones((256, 128), dtype=np.int64)
    g[:, 0] = 0
    g[:, -1] = 0
    yield "d2q9_channel_256x128", "D2Q9", g, "x", [("init_rho", np.ones(g.shape))], 1.0, [1e-6, 0.0], 5000, True
    g = load_geo("d3q19_pack_6x5x4_pxyz")
    yield "d3q19_pack_6x5x4_pxyz", "D3Q19", g, "xyz", [("init_rho", np.ones(g.shape))], 0.9, [1e-5, 2e-6, -1e-6], 30, False
    g = load_geo("d3q19_pack_6x5x4_px_walls")
    yield "d3q19_pack_6x5x4_px_walls", "D3Q19", g, "x", [("init_rho", np.ones(g.shape))], 1.0, [1e-5, 0.0, 0.0], 30, False
    g = load_geo("d3q19_open_4x4x3")
    yield "d3q19_open_4x4x3", "D3Q19", g, "", [("init_rho", np.ones(g.shape))], 0.8, [1e-5, 1e-6, 0.0], 10, False
    g, _ = sphere_geo((24, 24, 24), 4.0, 0.5, 7, "x", True)
    yield "d3q19_pack_24", "D3Q19", g, "x", [("init_rho", np.ones(g.shape))], 1.0, [1e-5, 0.0, 0.0], 200, False


def trt_cases():
    """Two-relaxation-time std_case (oracle/_ref/std_case_trt_*):
    (name, lattice, geo, periodic, attrs, (tau_sym, tau_anti), force, max_it)"""
    g = load_geo("d2q9_random_7x5_pxy")
    r = np.load(os.path.join(VT, "d2q9_random_7x5_pxy_attr_init_rho.npy"))
    yield "trt_d2q9_random_7x5", "D2Q9", g, "xy", [("init_rho", r)], (0.8, 1.3), [1e-5, -3e-6], 50
    g = np.ones((64, 32), dtype=np.int64)
    g[:, 0] = 0
    g[:, -1] = 0
    # (tau_sym - 1/2)(tau_anti - 1/2) = 3/16: the wall sits exactly half-way for Poiseuille flow
    yield "trt_d2q9_channel_64x32", "D2Q9", g, "x", [("init_rho", np.ones(g.shape))], (1.0, 0.875), [1e-5, 0.0], 500
    g = load_geo("d3q19_pack_6x5x4_px_walls")
    yield "trt_d3q19_pack_6x5x4_px_walls", "D3Q19", g, "x", [("init_rho", np.ones(g.shape))], (1.0, 0.6), \
        [1e-5, 0.0, 0.0], 30
    g = load_geo("d3q19_pack_6x5x4_pxyz")
    yield "trt_d3q19_pack_6x5x4_pxyz", "D3Q19", g, "xyz", [("init_rho", np.ones(g.shape))], (0.9, 1.7), \
        [1e-5, 2e-6, -1e-6], 30
    g, _ = sphere_geo((24, 24, 24), 4.0, 0.5, 7, "x", True)
    yield "trt_d3q19_pack_24", "D3Q19", g, "x", [("init_rho", np.ones(g.shape))], (0.7, 1.2), \
        [1e-5, 0.0, 0.0], 200


def carreau_table(n=40, lo=1e-16, hi=1e-4, nu0=1.0 / 6.0, nu_inf=0.02, lam2=1e10, power=0.5):
    """Viscosity against the strain-rate square sum_ij E_ij^2 (the table
    GeneralizedNewtonian reads, LBrheology.h:358-388): a shear-thinning
    Carreau curve sampled log-uniformly, printed with 17 digits."""
    sr = np.logspace(np.log10(lo), np.log10(hi), n)
    nu = nu_inf + (nu0 - nu_inf) * (1.0 + lam2 * sr) ** ((power - 1.0) / 2.0)
    text = "%d\n" % n + "".join("%.17g %.17g\n" % (a, b) for a, b in zip(sr, nu))
    table = np.array([[float(t) for t in line.split()] for line in text.splitlines()[1:]])
    return text, table


def rheo_cases():
    """polymer_viscosity_tabulation (oracle/_ref/rheology_*):
    (name, lattice, geo, periodic, attrs, force, max_it, table kwargs)"""
    g = np.ones((48, 24), dtype=np.int64)
    g[:, 0] = 0
    g[:, -1] = 0
    yield "rheo_d2q9_channel_48x24", "D2Q9", g, "x", [("init_rho", np.ones(g.shape))], [2e-5, 0.0], 400, {}
    g = load_geo("d2q9_random_7x5_pxy")
    r = np.load(os.path.join(VT, "d2q9_random_7x5_pxy_attr_init_rho.npy"))
    yield "rheo_d2q9_random_7x5", "D2Q9", g, "xy", [("init_rho", r)], [1e-4, -3e-5], 60, {"lam2": 1e8}
    g = load_geo("d3q19_pack_6x5x4_px_walls")
    yield "rheo_d3q19_pack_6x5x4_px_walls", "D3Q19", g, "x", [("init_rho", np.ones(g.shape))], [1e-4, 0.0, 0.0], \
        40, {"lam2": 1e9}
    g, _ = sphere_geo((20, 20, 20), 4.0, 0.5, 7, "x", True)
    yield "rheo_d3q19_pack_20", "D3Q19", g, "x", [("init_rho", np.ones(g.shape))], [5e-5, 1e-5, 0.0], 150, \
        {"lam2": 1e9}


D3Q19_C = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1],
                    [0, 1, -1], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
                    [-1, 0, 1], [0, -1, -1], [0, -1, 1]])


def one_phase_geometry(shape, seed, with_labels=True):
    """A std_one_phase input: pressure-ghost layers (tag 3, solid in the
    geometry) at z = 0 and z = nz-1, spherical obstacles (tag 0), fluid of
    phase 1 for x < nx/2 and phase 2 beyond, periodic in x and y.  Interface
    bits as std_one_phase reads them (main.cpp:266-273): 4 fluid-fluid,
    8 fluid-solid, 16 pressure boundary.  Interior domains 1 and 2 split the
    phase-2 fluid at z = nz/2; the force acts on phase 1 only."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    geo = np.ones(shape, dtype=np.int64)
    solid = np.zeros(shape, dtype=bool)
    for _ in range(3):
        c = rng.uniform([0, 0, 2], [nx, ny, nz - 2])
        r = rng.uniform(1.2, 2.2)
        solid |= (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < r * r
    ghost = (z == 0) | (z == nz - 1)
    geo[solid | ghost] = 0
    phase = np.where(x < nx // 2, 1, 2)
    base = np.where(geo == 1, phase, np.where(ghost, 3, 0))
    tags = base.copy()
    fluid = geo == 1
    for c in D3Q19_C:
        nbp = np.roll(base, shift=(-c[0], -c[1]), axis=(0, 1))
        nbp = np.roll(nbp, shift=-c[2], axis=2)  # z wrap only pairs ghost layers, never read for fluid
        other = (nbp == 3 - base) & (nbp >= 1) & (nbp <= 2)
        tags |= np.where(fluid & other, 4, 0)
        tags |= np.where(fluid & (nbp == 0), 8, 0)
        tags |= np.where(fluid & (nbp == 3), 16, 0)
    tags = np.where(fluid, tags, base)
    label = np.zeros(shape, dtype=np.int64)
    if with_labels:
        label[fluid & (phase == 2) & (z < nz // 2)] = 1
        label[fluid & (phase == 2) & (z >= nz // 2)] = 2
    force = np.where(fluid & (phase == 1), 1, 0).astype(np.int64)
    return geo, tags, label, force


def one_phase_cases():
    """(name, shape, seed, with_labels, tau, force, max_it, write)"""
    yield "op_tagged_10x8x12", (10, 8, 12), 5, True, 0.8, [0.0, 0.0, 1e-5], 100, 2
    yield "op_tagged_nolabel_9x7x10", (9, 7, 10), 8, False, 1.1, [2e-6, 0.0, 2e-5], 50, 1


def tp_cases():
    """(name, geo, periodic, rho0, wett, tau0, tau1, sigma, beta, momx, force, max_it)"""
    rng = np.random.default_rng(99)
    g = load_geo("d3q19_pack_6x5x4_pxyz")
    r0 = np.load(os.path.join(VT, "d3q19_pack_6x5x4_pxyz_attr_rho0.npy"))
    yield "tp_pack_6x5x4", g, "xyz", r0, np.ones(g.shape), 1.0, 1.0, 0.01, 0.7, 1e-4, [0.0, 0.0, 0.0], 30
    g, _ = sphere_geo((16, 16, 16), 3.0, 0.5, 11, "xyz", False)
    r0 = rng.choice([0.0, 0.25, 0.6, 1.0], size=g.shape)
    yield "tp_pack_16", g, "xyz", r0, np.full(g.shape, 0.8), 0.8, 1.2, 0.01, 0.7, 1e-4, [0.0, 1e-6, -2e-6], 100
    g, _ = sphere_geo((12, 14, 10), 3.0, 0.55, 5, "x", True)
    r0 = (rng.random(g.shape) > 0.5).astype(float)
    yield "tp_walls_12x14x10", g, "x", r0, np.full(g.shape, 1.0), 1.0, 1.0, 0.01, 0.7, 5e-5, [0.0, 0.0, 0.0], 60


def tp_d2q9_cases():
    """The twophase case built for D2Q9 (oracle/_ref/twophase_d2q9), same
    fields as tp_cases."""
    rng = np.random.default_rng(23)
    g = load_geo("d2q9_random_7x5_pxy")
    r0 = rng.choice([0.0, 0.4, 1.0], size=g.shape)
    yield "tp_d2q9_random_7x5", g, "xy", r0, np.ones(g.shape), 1.0, 1.0, 0.01, 0.7, 1e-4, [0.0, 1e-6], 40
    g = (rng.random((40, 18)) > 0.3).astype(np.int64)
    g[:, 0] = 0
    g[:, -1] = 0
    r0 = (rng.random(g.shape) > 0.5).astype(float)
    yield "tp_d2q9_walls_40x18", g, "x", r0, np.full(g.shape, 0.8), 0.8, 1.2, 0.01, 0.7, 5e-5, [0.0, -1e-6], 100


VTU_KEEP = {"d2q9_random_7x5", "d3q19_pack_6x5x4_pxyz", "trt_d3q19_pack_6x5x4_px_walls", "tp_pack_6x5x4"}


def main(selected):
    for name, lat, g, per, attrs, tau, force, max_it, prof in std_cases():
        if selected and name not in selected:
            continue
        binary = "std_case_d2q9" if lat == "D2Q9" else "std_case_d3q19"
        dat = input_dat(max_it, ["tau %s" % fmt(tau), "bodyforce %s" % fmt(force)])
        data = run_case(binary, g, lat, per, attrs, dat, keep=name if name in VTU_KEEP else None)
        params = dict(case="std_case", lattice=lat, periodic=per, tau=tau, force=force, steps=max_it)
        out = dict(geo=g.astype(np.int8), params=json.dumps(params),
                   init_rho=attrs[0][1].astype(np.float64))
        rho, vel = data["rho"], data["vel"]
        if prof:
            # the channel is x-invariant: keep one column (y = 1..ny-2) and the spread over x
            ny = g.shape[1] - 2
            out["rho_profile"] = rho.reshape(g.shape[0], ny)[0]
            out["velx_profile"] = vel[:, 0].reshape(g.shape[0], ny)[0]
            out["vely_profile"] = vel[:, 1].reshape(g.shape[0], ny)[0]
            out["x_spread"] = np.array([np.ptp(rho.reshape(g.shape[0], ny), axis=0).max(),
                                        np.ptp(vel[:, 0].reshape(g.shape[0], ny), axis=0).max()])
        else:
            out["rho"] = rho
            out["vel"] = vel
        np.savez_compressed(os.path.join(HERE, "ref_%s.npz" % name), **out)
        print("wrote ref_%s.npz" % name, rho.shape)
    for name, lat, g, per, attrs, taus, force, max_it in trt_cases():
        if selected and name not in selected:
            continue
        binary = "std_case_trt_d2q9" if lat == "D2Q9" else "std_case_trt_d3q19"
        dat = input_dat(max_it, ["tau %s" % fmt(list(taus)), "bodyforce %s" % fmt(force)])
        data = run_case(binary, g, lat, per, attrs, dat, keep=name if name in VTU_KEEP else None)
        params = dict(case="std_case_trt", lattice=lat, periodic=per, tau=list(taus), force=force, steps=max_it)
        np.savez_compressed(os.path.join(HERE, "ref_%s.npz" % name), geo=g.astype(np.int8),
                            params=json.dumps(params), init_rho=attrs[0][1].astype(np.float64),
                            rho=data["rho"], vel=data["vel"])
        print("wrote ref_%s.npz" % name, data["rho"].shape)
    for name, lat, g, per, attrs, force, max_it, tk in rheo_cases():
        if selected and name not in selected:
            continue
        text, table = carreau_table(**tk)
        binary = "rheology_d2q9" if lat == "D2Q9" else "rheology_d3q19"
        dat = input_dat(max_it, ["tau 1.0", "bodyforce %s" % fmt(force)])
        data = run_case(binary, g, lat, per, attrs, dat, extra_inputs={"test.dat": text})
        params = dict(case="rheology", lattice=lat, periodic=per, force=force, steps=max_it)
        np.savez_compressed(os.path.join(HERE, "ref_%s.npz" % name), geo=g.astype(np.int8),
                            params=json.dumps(params), init_rho=attrs[0][1].astype(np.float64), table=table,
                            rho=data["rho"], vel=data["vel"], viscosity=data["viscosity"])
        print("wrote ref_%s.npz" % name, data["rho"].shape, data["viscosity"].min(), data["viscosity"].max())
    for name, shape, seed, with_labels, tau, force, max_it, write in one_phase_cases():
        if selected and name not in selected:
            continue
        geo, tags, label, fon = one_phase_geometry(shape, seed, with_labels)
        attrs = [("nodetags", tags), ("domains", np.zeros(shape, dtype=np.int64)), ("force", fon),
                 ("interior_domains", label), ("normal_x", np.zeros(shape)), ("normal_y", np.zeros(shape)),
                 ("normal_z", np.zeros(shape))]
        dat = ("<names>\n    file tmp\n    directory run\n<end>\n"
               "<iterations>\n    max %d\n    write %d\n<end>\n<fluid>\n    tau %s\n    bodyforce %s\n<end>\n"
               % (max_it, write, fmt(tau), fmt(force)))
        data = run_one_phase(geo, attrs, dat)
        params = dict(case="std_one_phase", lattice="D3Q19", periodic="xy", tau=tau, force=force, steps=max_it,
                      rho_w=1.0)
        np.savez_compressed(os.path.join(HERE, "ref_%s.npz" % name), geo=geo.astype(np.int8), tags=tags,
                            label=label, force_on=fon, params=json.dumps(params), rho=data["rho"],
                            vel=data["vel"], flux=data["flux"])
        print("wrote ref_%s.npz" % name, data["rho"].shape, data["flux"])
    tp_all = [("D3Q19", "twophase") + c for c in tp_cases()] + [("D2Q9", "twophase_d2q9") + c for c in tp_d2q9_cases()]
    for lat, binary, name, g, per, r0, wett, tau0, tau1, sigma, beta, momx, force, max_it in tp_all:
        if selected and name not in selected:
            continue
        attrs = [("rho0", r0), ("rho1", 1.0 - r0), ("wettability", wett),
                 ("source", np.zeros(g.shape, dtype=np.int64))]
        dat = input_dat(max_it, ["tau %s" % fmt([tau0, tau1]), "bodyforce %s" % fmt(force),
                                 "sigma %s" % fmt(sigma), "beta %s" % fmt(beta), "momx %s" % fmt(momx)],
                        extra="outdir tp/\n")
        data = run_case(binary, g, lat, per, attrs, dat, twophase=True,
                        keep=name if name in VTU_KEEP else None)
        params = dict(case="twophase", lattice=lat, periodic=per, tau0=tau0, tau1=tau1, sigma=sigma,
                      beta=beta, momx=momx, force=force, steps=max_it)
        out = dict(geo=g.astype(np.int8), params=json.dumps(params), rho0_attr=r0, wett_attr=wett,
                   rho0=data["rho0"], rho1=data["rho1"], vel=data["vel"], force_x=data["force_x"])
        np.savez_compressed(os.path.join(HERE, "ref_%s.npz" % name), **out)
        print("wrote ref_%s.npz" % name, data["rho0"].shape, data["force_x"][-1])


if __name__ == "__main__":
    main(set(sys.argv[1:]))
