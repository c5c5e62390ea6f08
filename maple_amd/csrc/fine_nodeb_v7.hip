// Level-0 elastic operator variant 7: 2x2 nodes and 2 load cases per thread.
#include "fine_nodeb.h"

template void fine_apply_nodeb_g<double, 2, 2, 2>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply_nodeb_g<float, 2, 2, 2>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);
