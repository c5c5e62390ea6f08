// Host-side element matrices (row-major, FP64).
#pragma once

// Ke 24x24, Fe 24x6, X0 24x6 for a cube of side 1/r, isotropic (E, nu).
void hex8_elastic(int r, double E, double nu, double* Ke, double* Fe, double* X0);
// Ke 8x8, Fe 8x3, X0 8x3 for a cube of side 1/r, conductivity kappa.
void hex8_thermal(int r, double kappa, double* Ke, double* Fe, double* X0);
// G[c] = P_c^T Ke P_c for the 8 children c = cx*4 + cy*2 + cz of a coarse element.
void child_galerkin(int nd, const double* Ke, double* G);
