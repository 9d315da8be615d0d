// ops_stage.hip -- fused RK stage (placeholder translation unit).
#include "common.hpp"
