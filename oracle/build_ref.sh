#!/usr/bin/env bash
# Build the reference SIMP hot path from its own Fortran sources, where they
# lie, into oracle/_ref/ (objects, modules and the harness binary only).
#
#   SIMP_REF_DIR   reference checkout (default /root/reference)
#
# Recipe notes:
#  * modules are compiled in dependency order (globals, data_variables,
#    domain, boundary_domain, mpi_variables, tree_module, interfaces); the
#    reference Makefile's order (Makefile:114-121) does not build from clean.
#  * the golden-vector builds use no -fopenmp: without OpenMP the directives
#    are comments and the build is the serial reference, the same arithmetic
#    order as OMP_NUM_THREADS=1.  The omp_get_* calls that remain are
#    ordinary calls resolved by libomp at link time.  The OpenMP build
#    (omp_build below, timings only) rewrites the REDUCTION(-:...) clause at
#    single_step.f90:55, which flang rejects, on the way into the compiler.
#  * the shipped compiled module files (fortran/src/modules/*.mod) are never
#    put on the include path; modules are regenerated from source.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${SIMP_REF_DIR:-/root/reference}"
SRC="$REF/fortran/src"
OUT="$HERE/_ref"
FC="${FC:-/opt/rocm/bin/amdflang}"
FFLAGS="-O2 -ffp-contract=off -w"

if [ ! -d "$SRC" ]; then
    echo "build_ref: no reference sources at $SRC" >&2
    exit 3
fi
mkdir -p "$OUT/obj" "$OUT/mod"

mods="globals data_variables domain boundary_domain mpi_variables tree_module interfaces"
objs=()
for m in $mods; do
    "$FC" $FFLAGS -J"$OUT/mod" -c "$SRC/modules/$m.f90" -o "$OUT/obj/$m.o"
    objs+=("$OUT/obj/$m.o")
done
for d in ext_force in_out int_force integration neighbor phy_quantities sph; do
    for f in "$SRC/$d"/*.f90; do
        b="$(basename "$f" .f90)"
        "$FC" $FFLAGS -I"$OUT/mod" -J"$OUT/mod" -c "$f" -o "$OUT/obj/$b.o"
        objs+=("$OUT/obj/$b.o")
    done
done
"$FC" $FFLAGS -I"$OUT/mod" -J"$OUT/mod" -c "$HERE/ref_harness.f90" -o "$OUT/obj/ref_harness.o"
"$FC" $FFLAGS -o "$OUT/simp_ref_harness" "$OUT/obj/ref_harness.o" "${objs[@]}" \
    -L/opt/rocm/llvm/lib -lomp -Wl,-rpath,/opt/rocm/llvm/lib
echo "build_ref: $OUT/simp_ref_harness"
# the reference's CSV output routine (in_out/output.f90) behind a small
# driver, for the golden snapshots of the CSV writer
"$FC" $FFLAGS -I"$OUT/mod" -J"$OUT/mod" -c "$HERE/ref_output.f90" -o "$OUT/obj/ref_output.o"
"$FC" $FFLAGS -o "$OUT/simp_ref_output" "$OUT/obj/ref_output.o" "${objs[@]}" \
    -L/opt/rocm/llvm/lib -lomp -Wl,-rpath,/opt/rocm/llvm/lib
echo "build_ref: $OUT/simp_ref_output"
# the reference's initialize (scenario from its domain and boundary maps)
# behind a small driver, for the golden vectors of the scenario generator
"$FC" $FFLAGS -I"$OUT/mod" -J"$OUT/mod" -c "$HERE/ref_init.f90" -o "$OUT/obj/ref_init.o"
"$FC" $FFLAGS -o "$OUT/simp_ref_init" "$OUT/obj/ref_init.o" "${objs[@]}" \
    -L/opt/rocm/llvm/lib -lomp -Wl,-rpath,/opt/rocm/llvm/lib
echo "build_ref: $OUT/simp_ref_init"

# Switch variants.  Run parameters are compile-time PARAMETERs that the
# reference configures by editing globals.f90 (README.md:68-70); the edit is
# applied on the way into the compiler (stdin), so no copy of the source is
# written, and everything that uses the module is rebuilt against it.
variant() {   # <tag> (<PARAMETER text in globals.f90> <replacement>)...; sets vobjs
    local tag="$1" V="$OUT/$1"
    shift
    mkdir -p "$V/obj" "$V/mod"
    vobjs=()
    local edits=()
    while [ $# -ge 2 ]; do
        grep -q "$1" "$SRC/modules/globals.f90"   # the text the edit targets
        edits+=(-e "s/$1/$2/")
        shift 2
    done
    sed "${edits[@]}" "$SRC/modules/globals.f90" \
        | "$FC" $FFLAGS -x f95 -J"$V/mod" -c - -o "$V/obj/globals.o"
    vobjs+=("$V/obj/globals.o")
    for m in data_variables domain boundary_domain mpi_variables tree_module interfaces; do
        "$FC" $FFLAGS -I"$V/mod" -J"$V/mod" -c "$SRC/modules/$m.f90" -o "$V/obj/$m.o"
        vobjs+=("$V/obj/$m.o")
    done
    for d in ext_force in_out int_force integration neighbor phy_quantities sph; do
        for f in "$SRC/$d"/*.f90; do
            b="$(basename "$f" .f90)"
            "$FC" $FFLAGS -I"$V/mod" -J"$V/mod" -c "$f" -o "$V/obj/$b.o"
            vobjs+=("$V/obj/$b.o")
        done
    done
    "$FC" $FFLAGS -I"$V/mod" -J"$V/mod" -c "$HERE/ref_harness.f90" -o "$V/obj/ref_harness.o"
    "$FC" $FFLAGS -o "$OUT/simp_ref_harness_$tag" "$V/obj/ref_harness.o" "${vobjs[@]}" \
        -L/opt/rocm/llvm/lib -lomp -Wl,-rpath,/opt/rocm/llvm/lib
    echo "build_ref: $OUT/simp_ref_harness_$tag"
    eval "vobjs_$tag=(\"\${vobjs[@]}\")"
}
D1="$OUT/den1"
# summation density (den_alg = 1, density.f90:41-100)
variant den1 'parameter(den_alg = 3)' 'parameter(den_alg = 1)'
d1objs=("${vobjs[@]}")
# stress_2 particle approximation (pa_sph = 2, internal_stress.f90:144-216)
variant pa2 'parameter(pa_sph = 1)' 'parameter(pa_sph = 2)'
P2="$OUT/pa2"
p2objs=("${vobjs[@]}")
# Hibler and Gutfraind pressure (eos_alg = 1, 2; eos.f90:59, 94-102)
variant eos1 'parameter(eos_alg = 3)' 'parameter(eos_alg = 1)'
E1="$OUT/eos1"
e1objs=("${vobjs[@]}")
variant eos2 'parameter(eos_alg = 3)' 'parameter(eos_alg = 2)'
E2="$OUT/eos2"
e2objs=("${vobjs[@]}")
# Mohr-Coulomb rheology (rheology = 2; bulk_viscosity.f90:86-145,
# shear_viscosity.f90:67-124)
variant mc 'parameter(rheology = 1)' 'parameter(rheology = 2)'
MC="$OUT/mc"
mcobjs=("${vobjs[@]}")
# Wendland C2 and C4 kernels (skf = 4, 5; kernel.f90:359-464,
# boundary_kernel.f90:290-431)
variant c2 'parameter(skf = 6)' 'parameter(skf = 4)'
C2="$OUT/c2"
c2objs=("${vobjs[@]}")
variant c4 'parameter(skf = 6)' 'parameter(skf = 5)'
C4="$OUT/c4"
c4objs=("${vobjs[@]}")
# the quadratic, Gaussian, quintic-spline and Marquis kernels (skf = 1, 2,
# 3, 7, 8; kernel.f90:108-290, 527-633; boundary_kernel.f90:54-283)
for k in 1 2 3 7 8; do
    variant "k$k" 'parameter(skf = 6)' "parameter(skf = $k)"
    eval "kobjs_$k=(\"\${vobjs[@]}\")"
done
# Euler and leap-frog time integration (int_algorithm = 1, 3;
# time_integration.f90:86-148, 289-391)
variant euler 'parameter(int_algorithm = 2)' 'parameter(int_algorithm = 1)'
EU="$OUT/euler"
euobjs=("${vobjs[@]}")
variant leap 'parameter(int_algorithm = 2)' 'parameter(int_algorithm = 3)'
LF="$OUT/leap"
lfobjs=("${vobjs[@]}")
# continuity density (den_alg = 2; density.f90:107-165)
variant den2 'parameter(den_alg = 3)' 'parameter(den_alg = 2)'
D2="$OUT/den2"
d2objs=("${vobjs[@]}")
# Monaghan time step (tse = 1; time_step.f90:39-125)
variant tse1 'parameter(tse = 2)' 'parameter(tse = 1)'
T1="$OUT/tse1"
t1objs=("${vobjs[@]}")
# smoothing length kept (sle = 0) and Benz's rate (sle = 1;
# smoothing_length.f90:21-111), the latter also with the Euler scheme and
# with continuity density
variant sle0 'parameter(sle = 2)' 'parameter(sle = 0)'
variant sle1 'parameter(sle = 2)' 'parameter(sle = 1)'
variant sle1eu 'parameter(sle = 2)' 'parameter(sle = 1)' \
    'parameter(int_algorithm = 2)' 'parameter(int_algorithm = 1)'
variant sle1d2 'parameter(sle = 2)' 'parameter(sle = 1)' \
    'parameter(den_alg = 3)' 'parameter(den_alg = 2)'
# XSPH (regularization.f90:45-95; with it regularization also runs the
# artificial viscosity art_visc_alg selects, :29-35) and the artificial
# viscosities of Hernquist and Katz (:102-213) and Monaghan (:220-308),
# the latter also with the Monaghan time step, which then adds the
# artificial viscosity's CFL terms (time_step.f90:89-111)
variant xsph 'xsph_regularization = .FALSE.' 'xsph_regularization = .TRUE.'
variant avhk 'artificial_visc   = .FALSE.' 'artificial_visc   = .TRUE.'
variant avmon 'artificial_visc   = .FALSE.' 'artificial_visc   = .TRUE.' \
    'parameter(art_visc_alg = 1)' 'parameter(art_visc_alg = 2)'
variant avtse1 'artificial_visc   = .FALSE.' 'artificial_visc   = .TRUE.' \
    'parameter(tse = 2)' 'parameter(tse = 1)'
# the Wang and Marquis boundary forces (bfa = 2, 3; boundary_force.f90:109-224)
variant bfa2 'parameter(bfa = 1)' 'parameter(bfa = 2)'
variant bfa3 'parameter(bfa = 1)' 'parameter(bfa = 3)'
# boundary friction (boundary_force.f90:87-100 with the boundary kernels'
# gradients, boundary_kernel.f90; Wang's :157-166; Marquis' :217-221, which
# runs only with the boundary pressure off), the boundary pressure off
# (no boundary pairs at all, nnps.f90:440), and wind, water drag and the
# replacement pressure off (globals.f90:178-187)
FRIC='boundary_friction = .FALSE.'
FRIC_ON='boundary_friction = .TRUE.'
variant fric "$FRIC" "$FRIC_ON"
variant fric1 "$FRIC" "$FRIC_ON" 'parameter(skf = 6)' 'parameter(skf = 1)'
variant fricg "$FRIC" "$FRIC_ON" 'parameter(skf = 6)' 'parameter(skf = 2)'
variant fricq "$FRIC" "$FRIC_ON" 'parameter(skf = 6)' 'parameter(skf = 3)'
variant fric4 "$FRIC" "$FRIC_ON" 'parameter(skf = 6)' 'parameter(skf = 4)'
variant fric5 "$FRIC" "$FRIC_ON" 'parameter(skf = 6)' 'parameter(skf = 5)'
variant bfa2f "$FRIC" "$FRIC_ON" 'parameter(bfa = 1)' 'parameter(bfa = 2)'
variant bfa3f "$FRIC" "$FRIC_ON" 'parameter(bfa = 1)' 'parameter(bfa = 3)' \
    'boundary_pressure = .TRUE.' 'boundary_pressure = .FALSE.'
variant nobp 'boundary_pressure = .TRUE.' 'boundary_pressure = .FALSE.'
variant noforce 'wind_force        = .TRUE.' 'wind_force        = .FALSE.' \
    'water_force       = .TRUE.' 'water_force       = .FALSE.' \
    'pressure_force    = .TRUE.' 'pressure_force    = .FALSE.'
variant novisc 'viscosity         = .TRUE.' 'viscosity         = .FALSE.' \
    'pressure_force    = .TRUE.' 'pressure_force    = .FALSE.'
variant novisc1 'viscosity         = .TRUE.' 'viscosity         = .FALSE.' \
    'parameter(eos_alg = 3)' 'parameter(eos_alg = 1)'

# The OpenMP build, for the CPU baseline: the reference compiled as its
# Makefile intends (-fopenmp) and sized for the 250,000-particle workload.
# Two edits go in on the way into the compiler (stdin, as above; no source
# copy is written):
#  * maxn = 300000 (globals.f90:22; max_interaction = 40 maxn follows);
#  * single_step.f90:55 REDUCTION(-:...) written REDUCTION(+:...): flang
#    rejects the '-' reduction identifier (deprecated since OpenMP 5.2), and
#    OpenMP combines the private copies of a '-' reduction by addition, so
#    the two clauses are the same reduction.
# Timings only: with more than one thread the pair lists are merged in a
# thread-dependent order (nnps.f90:426-433), so no golden vector comes from
# this build.
omp_build() {
    local V="$OUT/omp" FO="-O2 -ffp-contract=off -w -fopenmp"
    mkdir -p "$V/obj" "$V/mod"
    local oobjs=()
    grep -q 'maxn = 80000' "$SRC/modules/globals.f90"
    grep -q 'REDUCTION(-:' "$SRC/integration/single_step.f90"
    sed 's/maxn = 80000/maxn = 300000/' "$SRC/modules/globals.f90" \
        | "$FC" $FO -x f95 -J"$V/mod" -c - -o "$V/obj/globals.o"
    oobjs+=("$V/obj/globals.o")
    for m in data_variables domain boundary_domain mpi_variables tree_module interfaces; do
        "$FC" $FO -I"$V/mod" -J"$V/mod" -c "$SRC/modules/$m.f90" -o "$V/obj/$m.o"
        oobjs+=("$V/obj/$m.o")
    done
    for d in ext_force in_out int_force integration neighbor phy_quantities sph; do
        for f in "$SRC/$d"/*.f90; do
            b="$(basename "$f" .f90)"
            if [ "$d/$b" = "integration/single_step" ]; then
                sed 's/REDUCTION(-:/REDUCTION(+:/' "$f" \
                    | "$FC" $FO -I"$V/mod" -J"$V/mod" -x f95 -c - -o "$V/obj/$b.o"
            else
                "$FC" $FO -I"$V/mod" -J"$V/mod" -c "$f" -o "$V/obj/$b.o"
            fi
            oobjs+=("$V/obj/$b.o")
        done
    done
    "$FC" $FO -I"$V/mod" -J"$V/mod" -c "$HERE/ref_harness.f90" -o "$V/obj/ref_harness.o"
    "$FC" $FO -o "$OUT/simp_ref_harness_omp" "$V/obj/ref_harness.o" "${oobjs[@]}" \
        -L/opt/rocm/llvm/lib -lomp -Wl,-rpath,/opt/rocm/llvm/lib
    echo "build_ref: $OUT/simp_ref_harness_omp"
}
omp_build
# the serial build with the same capacity, for the one-core figure at full size
variant big 'maxn = 80000' 'maxn = 300000'

# The drop-in check: the same harness and the reference's own
# time_integration, with single_step replaced by the iso_c_binding shim of
# simp_amd/fortran that calls libsimp_amd.so on the GPU; once per density
# variant (the shim compiled against that variant's modules).
LIB="$HERE/../simp_amd/lib/libsimp_amd.so"
gpu_harness() {   # <obj dir> <mod dir> <output> <objects...>
    local odir="$1" mdir="$2" out="$3"
    shift 3
    local FSRC="$HERE/../simp_amd/fortran"
    "$FC" $FFLAGS -I"$mdir" -J"$mdir" -c "$FSRC/simp_binding.f90" -o "$odir/simp_binding.o"
    "$FC" $FFLAGS -I"$mdir" -J"$mdir" -c "$FSRC/single_step_gpu.f90" -o "$odir/single_step_gpu.o"
    local gobjs=()
    for o in "$@"; do
        [ "$(basename "$o")" = "single_step.o" ] || gobjs+=("$o")
    done
    "$FC" $FFLAGS -o "$out" "$odir/ref_harness.o" "${gobjs[@]}" \
        "$odir/simp_binding.o" "$odir/single_step_gpu.o" \
        -L"$(dirname "$LIB")" -lsimp_amd -L/opt/rocm/llvm/lib -lomp \
        -Wl,-rpath,'$ORIGIN/../../simp_amd/lib' -Wl,-rpath,/opt/rocm/llvm/lib -Wl,-rpath,/opt/rocm/lib
    echo "build_ref: $out"
}
if [ -f "$LIB" ]; then
    gpu_harness "$OUT/obj" "$OUT/mod" "$OUT/simp_ref_harness_gpu" "${objs[@]}"
    gpu_harness "$D1/obj" "$D1/mod" "$OUT/simp_ref_harness_gpu_den1" "${d1objs[@]}"
    gpu_harness "$P2/obj" "$P2/mod" "$OUT/simp_ref_harness_gpu_pa2" "${p2objs[@]}"
    gpu_harness "$E1/obj" "$E1/mod" "$OUT/simp_ref_harness_gpu_eos1" "${e1objs[@]}"
    gpu_harness "$E2/obj" "$E2/mod" "$OUT/simp_ref_harness_gpu_eos2" "${e2objs[@]}"
    gpu_harness "$MC/obj" "$MC/mod" "$OUT/simp_ref_harness_gpu_mc" "${mcobjs[@]}"
    gpu_harness "$C2/obj" "$C2/mod" "$OUT/simp_ref_harness_gpu_c2" "${c2objs[@]}"
    gpu_harness "$C4/obj" "$C4/mod" "$OUT/simp_ref_harness_gpu_c4" "${c4objs[@]}"
    gpu_harness "$EU/obj" "$EU/mod" "$OUT/simp_ref_harness_gpu_euler" "${euobjs[@]}"
    gpu_harness "$LF/obj" "$LF/mod" "$OUT/simp_ref_harness_gpu_leap" "${lfobjs[@]}"
    gpu_harness "$T1/obj" "$T1/mod" "$OUT/simp_ref_harness_gpu_tse1" "${t1objs[@]}"
    gpu_harness "$D2/obj" "$D2/mod" "$OUT/simp_ref_harness_gpu_den2" "${d2objs[@]}"
    for k in 1 2 3 7 8; do
        eval "gpu_harness \"$OUT/k$k/obj\" \"$OUT/k$k/mod\" \"$OUT/simp_ref_harness_gpu_k$k\" \"\${kobjs_$k[@]}\""
    done
    for t in sle0 sle1 sle1eu sle1d2 xsph avhk avmon avtse1 bfa2 bfa3 fric fric1 fricg fricq \
             fric4 fric5 bfa2f bfa3f nobp noforce novisc novisc1; do
        eval "gpu_harness \"$OUT/$t/obj\" \"$OUT/$t/mod\" \"$OUT/simp_ref_harness_gpu_$t\" \"\${vobjs_$t[@]}\""
    done
fi
