// nn_nc5.hip — kernels instantiated for 5 components.
#include "kernels.hpp"

template struct nn::Launch<5>;
