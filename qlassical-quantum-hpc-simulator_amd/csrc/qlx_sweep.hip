// qlx_sweep.hip — windowed executor (LDS/register window sweeps).
#include "qlx_internal.hpp"

namespace qlx {

void execute_windowed(qlx_state*, const qlx_gate_instr*, size_t, const double*, size_t,
                      const qlx_matrix_entry*, size_t, const qlx_exec_opts&, qlx_exec_stats&) {
    fail(QLX_EUNSUPPORTED, "windowed execution is not built yet");
}

}  // namespace qlx
