// selbstkollision.cpp — restatement of selbstkollision.rs (filled in with the
// self-collision kernels). TEST INFRASTRUCTURE ONLY.
#include "gcoracle.h"
