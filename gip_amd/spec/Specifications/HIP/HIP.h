// The HIP specification (counterpart of Specifications/Multicore/Multicore.h
// and Specifications/Cuda/Cuda.h in the reference).
#include "GIP_ArchHIP.h"
#include "GIP_MatrixHIP.h"
#include "GIP_NonLinearHIP.h"
#include "GIP_TimerHIP.h"
#include "GIP_VectorHIP.h"
