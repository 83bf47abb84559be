// A CG solve under log::ProfilerHook (the profiling setup of the reference's
// examples/performance-debugging and the summary hooks of
// include/ginkgo/core/log/profiler_hook.hpp): a nested summary timed with
// HIP events on the executor's stream is printed as a Markdown table, and
// the same ranges go to roctx, where rocprofv3 --marker-trace records them
// next to the kernels.
//
// usage: profiled_solver <matrix.mtx> [max_iters=1000] [reduction=1e-10]
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "ginkgo_amd/ginkgo.hpp"

int main(int argc, char** argv)
{
    using vec = gko::matrix::Dense<double>;
    using cg = gko::solver::Cg<double>;
    using hook = gko::log::ProfilerHook;
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <matrix.mtx> [max_iters] [reduction]\n";
        return 1;
    }
    const gko::size_type max_iters = argc > 2 ? std::atoll(argv[2]) : 1000;
    const double reduction = argc > 3 ? std::atof(argv[3]) : 1e-10;

    auto exec = gko::Mi355xExecutor::create(0);
    {
        // the hooks are attached for the whole lifetime of the objects they
        // observe, so that every range they open is also closed
        auto summary = hook::create_nested_summary(
            gko::Timer::create_for_executor(exec),
            std::make_unique<hook::TableSummaryWriter>(std::cout, "CG solve on " +
                                                                      std::string(argv[1])));
        exec->add_logger(summary);
        std::shared_ptr<hook> roctx;
        try {
            roctx = hook::create_for_executor(exec);
            exec->add_logger(roctx);
        } catch (const gko::NotSupported& e) {
            std::cerr << "no roctx ranges: " << e.what() << "\n";
        }

        std::ifstream in(argv[1]);
        if (!in) {
            std::cerr << "cannot open " << argv[1] << "\n";
            return 1;
        }
        auto A = gko::share(gko::read_csr<double>(in, exec));
        summary->set_object_name(A.get(), "A");
        if (roctx) roctx->set_object_name(A.get(), "A");
        const auto n = A->get_size().rows;
        auto one = vec::create(exec, gko::dim<2>{n, 1});
        one->fill(1.0);
        auto b = vec::create(exec, gko::dim<2>{n, 1});
        auto x = vec::create(exec, gko::dim<2>{n, 1});
        A->apply(one.get(), b.get());
        x->fill(0.0);
        for (auto& h : {summary, roctx}) {
            if (!h) continue;
            h->set_object_name(b.get(), "b");
            h->set_object_name(x.get(), "x");
        }
        auto factory =
            cg::build()
                .with_criteria(gko::stop::Iteration::build().with_max_iters(max_iters),
                               gko::stop::ResidualNorm<double>::build().with_reduction_factor(
                                   reduction))
                .on(exec);
        auto solver = factory->generate(A);
        for (auto& h : {summary, roctx}) {
            if (h) h->set_object_name(solver.get(), "cg");
        }
        auto conv = gko::log::Convergence::create();
        solver->add_logger(conv);
        {
            auto range = summary->user_range("solve");
            solver->apply(b.get(), x.get());
        }
        exec->synchronize();
        std::cout << "iterations " << conv->get_num_iterations() << ", converged "
                  << (conv->has_converged() ? "yes" : "no") << "\n";
        exec->clear_loggers();
    }  // the summary hook is destroyed here and prints its table
    return 0;
}
