// Level-0 elastic operator variant 5: 2x1 nodes and 3 load cases per thread.
#include "fine_nodeb.h"

template void fine_apply_nodeb_g<double, 3, 2, 1>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply_nodeb_g<float, 3, 2, 1>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);
