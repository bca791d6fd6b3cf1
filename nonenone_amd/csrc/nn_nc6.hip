// nn_nc6.hip — kernels instantiated for 6 components.
#include "kernels.hpp"

template struct nn::Launch<6>;
