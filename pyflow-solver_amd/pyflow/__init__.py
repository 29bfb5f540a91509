# CFD Solver package root
