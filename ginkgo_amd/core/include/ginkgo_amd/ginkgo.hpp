// Umbrella header of the gko:: host library (the surface of
// include/ginkgo/ginkgo.hpp restricted to the sparse solver path).
#pragma once

#include "ginkgo_amd/base.hpp"
#include "ginkgo_amd/lin_op.hpp"
#include "ginkgo_amd/log.hpp"
#include "ginkgo_amd/profiler_hook.hpp"
#include "ginkgo_amd/timer.hpp"
#include "ginkgo_amd/matrix.hpp"
#include "ginkgo_amd/solver.hpp"
#include "ginkgo_amd/stop.hpp"
#include "ginkgo_amd/mtx_io.hpp"
#include "ginkgo_amd/batch.hpp"
