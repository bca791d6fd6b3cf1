// nn_nc8.hip — kernels instantiated for 8 components.
#include "kernels.hpp"

template struct nn::Launch<8>;
