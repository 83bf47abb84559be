// The fork's CG benchmark harness (examples/simple-solver/simple-solver.cpp of
// the reference) on the MI355X executor: reads a MatrixMarket file, sets
// b = A * 1 and x0 = 0, solves with CG (no preconditioner) to a relative
// residual of 1e-10 or 10000 iterations, runs one warm-up and one timed
// solve (wall clock around apply + synchronize), and appends
// `path,rows,nnz,time_ms,iters` to a CSV file.
//
// usage: simple_solver <matrix.mtx> [out.csv] [reduction=1e-10] [max_iters=10000]
#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "ginkgo_amd/ginkgo.hpp"

int main(int argc, char** argv)
{
    using vec = gko::matrix::Dense<double>;
    using mtx = gko::matrix::Csr<double, gko::int32>;
    using cg = gko::solver::Cg<double>;
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <matrix.mtx> [out.csv] [reduction] [max_iters]\n";
        return 1;
    }
    const std::string path = argv[1];
    const std::string csv = argc > 2 ? argv[2] : "ginkgo_cg_mi355x.csv";
    const double reduction = argc > 3 ? std::atof(argv[3]) : 1e-10;
    const gko::size_type max_iters = argc > 4 ? std::atoll(argv[4]) : 10000;

    auto exec = gko::Mi355xExecutor::create(0);
    std::ifstream in(path);
    if (!in) {
        std::cerr << "cannot open " << path << "\n";
        return 1;
    }
    auto A = gko::share(gko::read_csr<double>(in, exec));
    const auto n = A->get_size().rows;
    auto one = vec::create(exec, gko::dim<2>{n, 1});
    one->fill(1.0);
    auto b = vec::create(exec, gko::dim<2>{n, 1});
    A->apply(one.get(), b.get());

    auto solver = cg::build()
                      .with_criteria(gko::stop::Iteration::build().with_max_iters(max_iters),
                                     gko::stop::ResidualNorm<double>::build()
                                         .with_reduction_factor(reduction))
                      .on(exec)
                      ->generate(A);
    auto logger = gko::log::Convergence::create();
    solver->add_logger(logger);

    auto x = vec::create(exec, gko::dim<2>{n, 1});
    x->fill(0.0);
    solver->apply(b.get(), x.get());  // warm-up
    x->fill(0.0);
    exec->synchronize();
    const auto t0 = std::chrono::steady_clock::now();
    solver->apply(b.get(), x.get());
    exec->synchronize();
    const auto t1 = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    const auto iters = logger->get_num_iterations();
    std::cout << path << ": rows " << n << ", nnz " << A->get_num_stored_elements() << ", "
              << ms << " ms, " << iters << " iterations, residual "
              << (logger->get_residual_norm().empty() ? -1.0 : logger->get_residual_norm()[0])
              << "\n";
    std::ofstream out(csv, std::ios::app);
    out << path << "," << n << "," << A->get_num_stored_elements() << "," << ms << "," << iters
        << "\n";
    return 0;
}
