#!/bin/sh
# Builds oracle/_ref/ref_sk: the reference's Slater-Koster + strain bond
# routine (NEMO_3D/nemo3d/src/ham/mb_ham_spds_micro.c) with the nml vector
# and matrix sources it calls, compiled in place from /root/reference, and
# the harness oracle/ref_sk_driver.cpp.  Objects and the binary go to
# oracle/_ref/ only.  Skips quietly when the reference tree is absent.
set -e
R=/root/reference/NEMO_3D
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
[ -d "$R/nemo3d/src/ham" ] || { echo "ref_build: no reference tree, skipped"; exit 0; }
mkdir -p "$OUT/obj"
INC=""
for d in $(find "$R/nemo3d/src" -type d -not -path "*/CVS*" -not -path "*/html*" -not -path "*/latex*"); do
  INC="$INC -I$d"
done
X=$R/xerces-c-src2_1_0
INC="$INC -I$R/nml/include -I$R/eigen/include -I$R/XMLtoC/src -I$R/XMLtoC/apps/ioNemo3D/src -I$X/include -I$X/src -I$X/include/xercesc -I$X/src/xercesc"
FLAGS="-std=gnu++98 -fpermissive -w -DFAKE_MPI -UHamMemorySave -O1 -ffp-contract=off"
SRCS="nemo3d/src/ham/mb_ham_spds_micro.c
nml/src/vector/dvector/rvector.c
nml/src/vector/dvector/nml_dvector.c
nml/src/vector/ivector/ivector.c
nml/src/vector/ivector/nml_ivector.c
nml/src/vector/dcvector/cvector.c
nml/src/vector/dcvector/nml_dcvector.c
nml/src/scalar/dcscalar/complex.c
nml/src/global/nml_memory.c
nml/src/matrix/dcmatrix/cmatrix.c
nml/src/matrix/dcmatrix/nml_dcmatrix.c
nml/src/matrix/dmatrix/rmatrix.c
nml/src/matrix/dmatrix/nml_dmatrix.c
nml/src/scalar/dscalar/nml_dscalar.c
nml/src/scalar/dcscalar/nml_dcscalar.c
nml/src/scalar/iscalar/nml_iscalar.c
xerces-c-src2_1_0/src/xercesc/framework/XMLEntityDecl.cpp"
OBJS=""
for f in $SRCS; do
  o="$OUT/obj/$(basename "$f" | sed 's/\.[a-zA-Z]*$//').o"
  g++ $FLAGS $INC -c "$R/$f" -o "$o"
  OBJS="$OBJS $o"
done
g++ $FLAGS $INC -c "$HERE/ref_sk_driver.cpp" -o "$OUT/obj/ref_sk_driver.o"
# functions of the compiled units that the bond routine never calls stay
# unresolved and are bound lazily (never, since they are not called)
g++ -no-pie -Wl,-z,lazy -Wl,--unresolved-symbols=ignore-all -o "$OUT/ref_sk" \
  "$OUT/obj/ref_sk_driver.o" $OBJS -lm
echo "ref_build: built $OUT/ref_sk"
