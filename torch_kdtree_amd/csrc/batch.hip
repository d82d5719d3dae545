// Batched many-cloud mode (placeholder translation unit, filled in below).
#include "knn_gpu.hpp"
