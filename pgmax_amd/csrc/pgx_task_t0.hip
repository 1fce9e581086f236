// Task-loop kernels for max-product (T = 0).
#include "pgx_task.h"

namespace pgx {
template int launch_task_loop<true>(pgx_graph*, const RunBufs&, int32_t, int, hipStream_t);
}  // namespace pgx
