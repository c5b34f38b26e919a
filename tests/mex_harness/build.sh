#!/bin/bash
# Builds the MEX gateways against the minimal mxArray layer of this folder
# and links them with libdugma_hip.so into one test driver: ./harness
set -e
H=$(cd "$(dirname "$0")" && pwd)
R=$(cd "$H/../.." && pwd)
G=$R/dugma_amd/mex
O=${1:-$H/harness}
CXX=${CXX:-g++}
FL="-std=c++17 -O1 -I$H"
$CXX $FL -DmexFunction=mex_init -c "$G/init.cpp" -o "$O.init.o"
$CXX $FL -DmexFunction=mex_update_sig -c "$G/update_sig.cpp" -o "$O.update_sig.o"
$CXX $FL -DmexFunction=mex_compi -c "$G/compi.cpp" -o "$O.compi.o"
$CXX $FL -DmexFunction=mex_fr_m -c "$G/fr_m.cpp" -o "$O.fr_m.o"
$CXX $FL -DmexFunction=mex_energy_grad -c "$G/energy_grad.cpp" -o "$O.energy_grad.o"
$CXX $FL "$H/harness.cpp" "$O.init.o" "$O.update_sig.o" "$O.compi.o" "$O.fr_m.o" "$O.energy_grad.o" -o "$O" \
    -L"$R/dugma_amd" -ldugma_hip -Wl,-rpath,"$R/dugma_amd"
rm -f "$O".*.o
