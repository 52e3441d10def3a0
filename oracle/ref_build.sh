#!/bin/bash
# Builds oracle/_ref/iroiro_ref: the reference IroIro++ Wilson solve path,
# compiled from its own sources where they lie under the reference tree,
# linked with oracle/ref_driver.cpp (a command-line driver).  Test
# infrastructure only: used to write golden fixtures and to time the
# reference CPU path.  Outputs go to oracle/_ref/ only.
#
# Flags follow the reference's configure defaults for a single-process,
# non-BG/Q build with the improved Wilson kernel (configure.ac:88-95).
# domainWallCore.cpp is listed only because lib/Fopr/fopr.h pulls in
# Dirac_DomainWall, whose inline virtuals reference it.
set -euo pipefail
REF=${IROIRO_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/lib/Dirac_ops" ]; then
  echo "ref_build: reference tree not found at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
CXXFLAGS="-std=gnu++98 -O3 -fopenmp -w -D_GLIBCXX_USE_CXX11_ABI=0 -DIMPROVED_WILSON -DVERBOSITY=1"
INC="-I$REF -I$REF/include -I$REF/lib"
SRCS="
lib/Dirac_ops/dirac_wilson.cpp
lib/Dirac_ops/dirac.cpp
lib/Dirac_ops/wilsonLikeUtils.cpp
lib/Dirac_ops/eoUtils.cpp
lib/Dirac_ops/domainWallCore.cpp
lib/Solver/solver.cpp
lib/Solver/solver_CG.cpp
lib/Solver/multiShiftSolver_CG.cpp
lib/Geometry/siteIndex.cpp
lib/Geometry/siteIndex_EvenOdd.cpp
lib/Geometry/siteMap.cpp
lib/Communicator/communicator_single.cpp
lib/Communicator/comm_io.cpp
lib/Main/commonPrms.cpp
lib/Tools/ErrorsHandling/base_errors.cpp
lib/Tools/ErrorsHandling/convergence_errors.cpp
lib/PugiXML/pugixml.cpp
lib/PugiXML/pugi_interface.cpp
lib/Tools/fieldUtils.cpp
lib/Tools/sunMatUtils.cpp
lib/Geometry/shiftField.cpp
lib/Geometry/autoMap_improved.cpp
lib/IO/binary_reader.cpp
lib/IO/binary_writer.cpp
lib/IO/NERSC_writer.cpp
lib/Measurements/GaugeM/staples.cpp
lib/Tools/gammaMatrices.cpp
"
OBJS=""
for s in $SRCS; do
  o="$OUT/obj/$(echo "$s" | tr '/' '_' | sed 's/\.cpp$/.o/')"
  FLAGS="$CXXFLAGS"
  # The IO classes' header()/check() are non-void functions without a
  # return statement; g++ 8+ turns flowing off their end into a trap at -O1
  # and above, so these translation units are compiled at -O0.
  case "$s" in lib/IO/*) FLAGS="$(echo "$CXXFLAGS" | sed 's/-O3/-O0/')" ;; esac
  if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ] || [ "$0" -nt "$o" ]; then
    g++ $FLAGS $INC -c "$REF/$s" -o "$o"
  fi
  OBJS="$OBJS $o"
done
g++ $CXXFLAGS $INC -c "$HERE/ref_driver.cpp" -o "$OUT/obj/ref_driver.o"
g++ -fopenmp -o "$OUT/iroiro_ref" "$OUT/obj/ref_driver.o" $OBJS
echo "ref_build: built $OUT/iroiro_ref"

# Drop-in test binary: the reference's own Inverter_WilsonLike / Qprop_EvenOdd
# driving the device classes of iroiro_amd/host/iroiro_ref_adapter.hpp
# (needs the HIP library; links it with an rpath relative to this folder).
REPO=$(cd "$HERE/.." && pwd)
HIPLIB="$REPO/iroiro_amd/libiroiro_hip.so"
if [ -f "$HIPLIB" ]; then
  QOBJ=""
  for QSRC in lib/Measurements/FermionicM/qprop_EvenOdd.cpp \
              lib/Measurements/FermionicM/qprop_MultiShift.cpp; do
    o="$OUT/obj/$(echo "$QSRC" | tr '/' '_' | sed 's/\.cpp$/.o/')"
    if [ ! -f "$o" ] || [ "$REF/$QSRC" -nt "$o" ]; then
      g++ $CXXFLAGS $INC -c "$REF/$QSRC" -o "$o"
    fi
    QOBJ="$QOBJ $o"
  done
  g++ $CXXFLAGS $INC -I"$REPO/include" -I"$REPO/iroiro_amd/host" \
      -c "$REPO/tests/cpp/ref_adapter_driver.cpp" -o "$OUT/obj/ref_adapter_driver.o"
  g++ -fopenmp -o "$OUT/iroiro_ref_hip" "$OUT/obj/ref_adapter_driver.o" $OBJS $QOBJ \
      -L"$REPO/iroiro_amd" -liroiro_hip -Wl,-rpath,'$ORIGIN/../../iroiro_amd'
  echo "ref_build: built $OUT/iroiro_ref_hip"
fi
