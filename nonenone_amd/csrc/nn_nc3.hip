// nn_nc3.hip — kernels instantiated for 3 components.
#include "kernels.hpp"

template struct nn::Launch<3>;
