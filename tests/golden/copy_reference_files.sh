#!/bin/sh
# Copies the reference's shipped run-configuration files (data, not source) into
# tests/golden/reference_files/ as fixtures for the parameter-reader and
# field-record tests.  Run from the repository root in a checkout that has the
# reference at REF (default /root/reference).
REF=${REF:-/root/reference}
D=tests/golden/reference_files
mkdir -p $D
cp $REF/3D_NavierStokes/bin/params.txt      $D/ns3d_params.txt
cp $REF/3D_NavierStokes/bin/paramsEXTRA.txt $D/ns3d_paramsEXTRA.txt
cp $REF/3D_Boussinesq/bin/params.txt        $D/bous_params.txt
cp $REF/3D_Boussinesq/bin/paramsEXTRA.txt   $D/bous_paramsEXTRA.txt
cp $REF/3D_Euler/bin/params.txt             $D/euler_params.txt
cp $REF/3D_Euler/bin/Zk.out                 $D/euler_Zk.out
chmod u+w $D/*
