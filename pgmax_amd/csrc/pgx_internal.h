// Declarations shared by the library's translation units (not part of the C ABI).
#ifndef PGX_INTERNAL_H_
#define PGX_INTERNAL_H_

#include <string>

#include "../../include/pgx.h"
#include "pgx_plan.h"

namespace pgx {

// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string& msg);

// The compiled plan of a graph.
const Plan& graph_plan(const pgx_graph* g);

}  // namespace pgx

#endif  // PGX_INTERNAL_H_
