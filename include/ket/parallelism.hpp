// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/parallelism.hpp).
#pragma once

namespace ket {

// Host threads for CPU-side work. The GPU backend evolves states on the
// device, so the count is recorded and reported but no kernel depends on it.
// KET_NUM_THREADS sets the initial value; 0 selects the hardware concurrency.
void set_num_threads(unsigned threads);
unsigned num_threads();

}  // namespace ket
