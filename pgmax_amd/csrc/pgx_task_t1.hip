// Task-loop kernels for T > 0.
#include "pgx_task.h"

namespace pgx {
template int launch_task_loop<false>(pgx_graph*, const RunBufs&, int32_t, int, hipStream_t);
}  // namespace pgx
