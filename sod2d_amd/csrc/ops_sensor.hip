// ops_sensor.hip -- entropy-viscosity sensor (placeholder translation unit).
#include "common.hpp"
