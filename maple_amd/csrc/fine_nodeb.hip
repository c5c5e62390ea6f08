// Level-0 elastic operator, node blocks with 2x1 / 2x2 nodes per thread (variants 4-7).
#include "kernels.h"

template <typename T, int NCG, int NX, int NZ>
void fine_apply_nodeb_g(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u, const T* f,
                        const T* dinv, T omega, T* out, double* partial, hipStream_t s);

// variants 4, 5: 2x1 nodes, six / three cases per thread; 6, 7: 2x2 nodes, three / two
template <typename T>
void fine_apply_nodeb(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u, const T* f,
                      const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  switch (L.var) {
    case 4: fine_apply_nodeb_g<T, 6, 2, 1>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s); break;
    case 5: fine_apply_nodeb_g<T, 3, 2, 1>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s); break;
    case 6: fine_apply_nodeb_g<T, 3, 2, 2>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s); break;
    default: fine_apply_nodeb_g<T, 2, 2, 2>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s); break;
  }
}
template void fine_apply_nodeb<double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply_nodeb<float>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);

