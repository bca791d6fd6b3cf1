// nn_nc4.hip — kernels instantiated for 4 components.
#include "kernels.hpp"

template struct nn::Launch<4>;
