#!/usr/bin/env bash
# Build the reference BADChIMP cases used as the parity checker, from the
# sources where they lie under /root/reference, into oracle/_ref/ only.
#
#   std_case_d2q9   src/std_case/main.cpp as shipped (#define LT D2Q9)
#   std_case_d3q19  the same file with its lattice line switched to D3Q19
#                   (std_case/main.cpp:14-15 offers the switch); the edited
#                   copy is written to oracle/_ref/ and never committed
#   twophase        src/twophase/main_TWOPHASE.cpp as shipped (D3Q19)
#   twophase_d2q9   the same file with its lattice lines switched to D2Q9
#                   (main_TWOPHASE.cpp:50-51 offers the switch)
#   std_case_trt_d2q9, std_case_trt_d3q19
#                   std_case/main.cpp with its collision switched to the
#                   reference's two-relaxation-time operator: calcOmegaBGK ->
#                   calcOmegaBGKTRT (LBcollision.h:50-75) and calcDeltaOmegaF
#                   -> calcDeltaOmegaFTRT with phi = 1 (LBcollision.h:214-233);
#                   tau = fluid.tau[0] (symmetric), tauAnti = fluid.tau[1]
#   rheology_d2q9, rheology_d3q19
#                   src/polymer_viscosity_tabulation/main.cpp (the std_case
#                   loop with GeneralizedNewtonian, LBrheology.h:321-476)
#                   with its hard-coded input directory switched to ./../
#                   (the std_case layout), its body force read the way
#                   std_case/main.cpp:59 reads it (the shipped
#                   `bodyForce.set(0, 0) = input[...]` line does not compile
#                   against the current Input class: ambiguous operator=),
#                   and, for D3Q19, its lattice line
#   one_phase       src/std_one_phase/main.cpp (tagged-link boundaries and
#                   per-domain mass sources, D3Q19) with its hard-coded
#                   input directory switched to ./../
#   flux_ref        oracle/flux_harness.cpp: the reference's
#                   calcCapNumbForceCartDir and calcFluxForceCartDir
#                   (LBglobalforcing.h:8-98) on fields read from a file
#   init_ref        oracle/init_harness.cpp: the reference's initiateLbField
#                   (LBinitiatefield.h:33-57) on fields read from a file
#   std_case_gpu_d2q9, std_case_gpu_d3q19, twophase_gpu_d2q9, twophase_gpu_d3q19
#                   tests/integration/{std_case,twophase}_gpu.cpp: the
#                   reference's set-up classes driving this repository's
#                   library through the C++ shim (needs
#                   badchimp-cpp_amd/lib/libbadchimp_amd.so)
#
# Flags follow the reference release build (CMakeLists.txt:71): -O3 -fopenmp
# -std=c++17.  MPI is MPICH from /opt/conda (the reference links MPI via
# find_package).  At run time only libmpi.so.12 goes on LD_LIBRARY_PATH
# (oracle/_ref/mpilib), because the rest of /opt/conda/lib carries an old
# libstdc++ (with the Fortran runtime libmpi itself needs).
#
# Needs /root/reference; exits 0 without building when it is absent.
set -euo pipefail
REF=/root/reference/src
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF" ]; then
  echo "build_ref: /root/reference absent, nothing to build"
  exit 0
fi
if [ ! -e /opt/conda/lib/libmpi.so ]; then
  echo "build_ref: MPICH (/opt/conda/lib/libmpi.so) absent, reference unbuildable"
  exit 0
fi
mkdir -p "$OUT/mpilib"
for lib in libmpi.so.12 libgfortran.so.4 libquadmath.so.0; do
  ln -sf "/opt/conda/lib/$lib" "$OUT/mpilib/$lib"
done
CXXFLAGS="-std=c++17 -O3 -fopenmp -w -I/opt/conda/include"
LIBS="$REF/lbsolver/LBd2q9.cpp $REF/lbsolver/LBd3q19.cpp /opt/conda/lib/libmpi.so"

sed 's/^#define LT D2Q9/#define LT D3Q19/' "$REF/std_case/main.cpp" > "$OUT/std_case_d3q19.cpp.new"
if cmp -s "$OUT/std_case_d3q19.cpp.new" "$OUT/std_case_d3q19.cpp"; then
  rm "$OUT/std_case_d3q19.cpp.new"
else
  mv "$OUT/std_case_d3q19.cpp.new" "$OUT/std_case_d3q19.cpp"
fi

# the two-relaxation-time variants (edited copies, never committed)
trt_edit() {  # lattice line
  sed -e "s/^#define LT D2Q9/#define LT $1/" \
      -e 's/lbBase_t tau = input\["fluid"\]\["tau"\];/lbBase_t tau = input["fluid"]["tau"][0]; lbBase_t tauAnti = input["fluid"]["tau"][1];/' \
      -e 's/calcOmegaBGK<LT>(fNode, tau, rhoNode, u2, cu)/calcOmegaBGKTRT<LT>(fNode, tau, tauAnti, rhoNode, u2, cu)/' \
      -e 's/calcDeltaOmegaF<LT>(tau, cu, uF, cF)/calcDeltaOmegaFTRT<LT>(tau, tauAnti, 1.0, cu, uF, cF)/' \
      "$REF/std_case/main.cpp"
}
for lat in D2Q9 D3Q19; do
  f="$OUT/std_case_trt_$(echo $lat | tr 'A-Z' 'a-z').cpp"
  trt_edit $lat > "$f.new"
  if ! grep -q calcOmegaBGKTRT "$f.new" || ! grep -q calcDeltaOmegaFTRT "$f.new" || ! grep -q tauAnti "$f.new"; then
    echo "build_ref: the TRT switch did not apply to std_case/main.cpp" >&2
    exit 1
  fi
  if cmp -s "$f.new" "$f"; then rm "$f.new"; else mv "$f.new" "$f"; fi
done

# the generalized-Newtonian case with a relative input directory
for lat in D2Q9 D3Q19; do
  f="$OUT/rheology_$(echo $lat | tr 'A-Z' 'a-z').cpp"
  sed -e "s/^#define LT D2Q9/#define LT $lat/" \
      -e 's|std::string chimpDir = "/home/[^"]*";|std::string chimpDir = "./../";|' \
      -e '/^    VectorField<LT> bodyForce(1, 1);$/d' \
      -e 's|^    bodyForce.set(0, 0) = input\["fluid"\]\["bodyforce"\];|    VectorField<LT> bodyForce(1, 1, input["fluid"]["bodyforce"]);|' \
      "$REF/polymer_viscosity_tabulation/main.cpp" > "$f.new"
  if ! grep -q 'chimpDir = "./../"' "$f.new" || ! grep -q "GeneralizedNewtonian" "$f.new" ||
     ! grep -q 'VectorField<LT> bodyForce(1, 1, input' "$f.new"; then
    echo "build_ref: the input-directory switch did not apply to polymer_viscosity_tabulation/main.cpp" >&2
    exit 1
  fi
  if cmp -s "$f.new" "$f"; then rm "$f.new"; else mv "$f.new" "$f"; fi
done

f="$OUT/twophase_d2q9.cpp"
sed -e 's|^//#define LT D2Q9|#define LT D2Q9|' -e 's|^#define LT D3Q19|//#define LT D3Q19|' \
    "$REF/twophase/main_TWOPHASE.cpp" > "$f.new"
if ! grep -q '^#define LT D2Q9' "$f.new" || grep -q '^#define LT D3Q19' "$f.new"; then
  echo "build_ref: the lattice switch did not apply to twophase/main_TWOPHASE.cpp" >&2
  exit 1
fi
if cmp -s "$f.new" "$f"; then rm "$f.new"; else mv "$f.new" "$f"; fi

f="$OUT/one_phase.cpp"
sed -e 's|std::string chimpDir = "/home/[^"]*";|std::string chimpDir = "./../";|' \
    "$REF/std_one_phase/main.cpp" > "$f.new"
if ! grep -q 'chimpDir = "./../"' "$f.new"; then
  echo "build_ref: the input-directory switch did not apply to std_one_phase/main.cpp" >&2
  exit 1
fi
if cmp -s "$f.new" "$f"; then rm "$f.new"; else mv "$f.new" "$f"; fi

build() {  # name, main source, extra include dir
  local name=$1 src=$2 inc=$3
  if [ ! -x "$OUT/$name" ] || [ "$src" -nt "$OUT/$name" ]; then
    g++ $CXXFLAGS -I"$inc" "$src" $LIBS -o "$OUT/$name"
    echo "build_ref: built $name"
  fi
}
build std_case_d2q9 "$REF/std_case/main.cpp" "$REF/std_case"
build std_case_d3q19 "$OUT/std_case_d3q19.cpp" "$REF/std_case"
build twophase "$REF/twophase/main_TWOPHASE.cpp" "$REF/twophase"
build twophase_d2q9 "$OUT/twophase_d2q9.cpp" "$REF/twophase"
build std_case_trt_d2q9 "$OUT/std_case_trt_d2q9.cpp" "$REF/std_case"
build std_case_trt_d3q19 "$OUT/std_case_trt_d3q19.cpp" "$REF/std_case"
build rheology_d2q9 "$OUT/rheology_d2q9.cpp" "$REF/polymer_viscosity_tabulation"
build rheology_d3q19 "$OUT/rheology_d3q19.cpp" "$REF/polymer_viscosity_tabulation"
build one_phase "$OUT/one_phase.cpp" "$REF/std_one_phase"
# the reference's global-forcing functions on given fields (our own harness
# source, compiled against the reference's headers)
if [ ! -x "$OUT/flux_ref" ] || [ "$HERE/flux_harness.cpp" -nt "$OUT/flux_ref" ]; then
  g++ $CXXFLAGS -I"$REF/lbsolver" "$HERE/flux_harness.cpp" $LIBS -o "$OUT/flux_ref"
  echo "build_ref: built flux_ref"
fi
if [ ! -x "$OUT/init_ref" ] || [ "$HERE/init_harness.cpp" -nt "$OUT/init_ref" ]; then
  g++ $CXXFLAGS -I"$REF/lbsolver" "$HERE/init_harness.cpp" $LIBS -o "$OUT/init_ref"
  echo "build_ref: built init_ref"
fi

# the drop-in demonstration (test infrastructure): reference set-up + shim
REPO="$(cd "$HERE/.." && pwd)"
LIBDIR="$REPO/badchimp-cpp_amd/lib"
DEMO="$REPO/tests/integration/std_case_gpu.cpp"
SHIM="$REPO/badchimp-cpp_amd/include/lbamd/lbsolver_gpu.hpp"
if [ -e "$LIBDIR/libbadchimp_amd.so" ]; then
  for lat in D2Q9 D3Q19; do
    name=std_case_gpu_$(echo $lat | tr 'A-Z' 'a-z')
    if [ ! -x "$OUT/$name" ] || [ "$DEMO" -nt "$OUT/$name" ] || [ "$SHIM" -nt "$OUT/$name" ]; then
      g++ -std=c++17 -O2 -w -I/opt/conda/include -I"$REF" -I"$REPO/badchimp-cpp_amd/include" -DLB_LATTICE=$lat \
        "$DEMO" "$REF/lbsolver/LBd2q9.cpp" "$REF/lbsolver/LBd3q19.cpp" /opt/conda/lib/libmpi.so \
        -L"$LIBDIR" -lbadchimp_amd -Wl,-rpath,'$ORIGIN/../../badchimp-cpp_amd/lib' -o "$OUT/$name"
      echo "build_ref: built $name"
    fi
  done
  TPDEMO="$REPO/tests/integration/twophase_gpu.cpp"
  for lat in D2Q9 D3Q19; do
    name=twophase_gpu_$(echo $lat | tr 'A-Z' 'a-z')
    if [ ! -x "$OUT/$name" ] || [ "$TPDEMO" -nt "$OUT/$name" ] || [ "$SHIM" -nt "$OUT/$name" ]; then
      g++ -std=c++17 -O2 -w -I/opt/conda/include -I"$REF" -I"$REPO/badchimp-cpp_amd/include" -DLB_LATTICE=$lat \
        "$TPDEMO" "$REF/lbsolver/LBd2q9.cpp" "$REF/lbsolver/LBd3q19.cpp" /opt/conda/lib/libmpi.so \
        -L"$LIBDIR" -lbadchimp_amd -Wl,-rpath,'$ORIGIN/../../badchimp-cpp_amd/lib' -o "$OUT/$name"
      echo "build_ref: built $name"
    fi
  done
fi
