#!/bin/bash
# Clears old outputs and runs the program on Input.txt in this folder
# (the steps of the reference run script; build first with __graft_entry__.build()).
cd "$(dirname "$0")" || exit 1
rm -f ./*.out ./*.vtk
exec ../../dba_amd/bin/KH_DBA
