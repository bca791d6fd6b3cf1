// nn_nc2.hip — kernels instantiated for 2 components.
#include "kernels.hpp"

template struct nn::Launch<2>;
