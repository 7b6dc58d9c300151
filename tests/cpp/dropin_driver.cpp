// dropin_driver.cpp -- exercises the C++ drop-in classes of prispack_amd.hpp
// the way PrisPack's own test drivers use its host classes
// (source/test/*: build A, make a preconditioner, setup(), solve(A, b, x),
// read getNumIter/getError/getConvergence).
//
// usage: dropin_driver <input.bin> <output.bin> <cg|gmres> <ict|ilut|ildlt|ic0|ilu0|none> <tau> <tol>
// input.bin : u32 n, u32 nnz, u32 rowptr[n+1], u32 col[nnz], f64 val[nnz], f64 b[n]
// output.bin: f64 x_host[n] (host drop-in solve), f64 x_dev[n] (device-resident solve)
// stdout    : one JSON line with the getters and the interface checks.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

#include "prispack_amd.hpp"

using namespace prispack;

namespace {

template <typename T>
void read(std::ifstream &f, T *p, size_t n) {
    f.read(reinterpret_cast<char *>(p), (std::streamsize)(sizeof(T) * n));
    if (!f) throw std::runtime_error("short input file");
}

double maxdiff(const HostVector<double> &a, const HostVector<double> &b, uint32_t off_b = 0) {
    double m = 0.0;
    for (uint32_t i = 0; i < a.getLength(); ++i) m = std::fmax(m, std::fabs(a[i] - b[i + off_b]));
    return m;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s in out cg|gmres ict|ilut|ildlt|ic0|ilu0|none tau tol\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream f(argv[1], std::ios::binary);
        uint32_t n = 0, nnz = 0;
        read(f, &n, 1);
        read(f, &nnz, 1);
        std::vector<uint32_t> rp(n + 1), ci(nnz);
        std::vector<double> v(nnz), bv(n);
        read(f, rp.data(), n + 1);
        read(f, ci.data(), nnz);
        read(f, v.data(), nnz);
        read(f, bv.data(), n);
        const std::string solver = argv[3], pc = argv[4];
        const double tau = std::atof(argv[5]), tol = std::atof(argv[6]);

        CSRMatrix<double> A(n, n, rp, ci, v);
        auto dA = std::make_shared<DeviceCSRMatrix>(A);
        std::shared_ptr<DevicePrecondition> M;
        if (pc == "ict") M = std::make_shared<DeviceIncompleteCholesky>(dA, tau);
        else if (pc == "ilut") M = std::make_shared<DeviceIncompleteLU>(dA, 20, tau);
        else if (pc == "ildlt") M = std::make_shared<DeviceIncompleteLDLT>(dA, tau);
        else if (pc == "ic0") M = std::make_shared<DeviceIC0>(dA);
        else if (pc == "ilu0") M = std::make_shared<DeviceILU0>(dA);
        else if (pc != "none") throw std::runtime_error("unknown preconditioner " + pc);

        // interface checks on the preconditioner
        double d_host_dev = -1, d_offset = -1, d_inout = -1, d_parts = -1, fill = 0;
        int ready_before = -1, ready_after = -1;
        if (M) {
            ready_before = M->isReady();
            M->setup();
            ready_after = M->isReady();
            fill = M->getPreconditionFillinRatio();
            HostVector<double> xh(bv);
            M->MInvSolve(xh);  // host vector, in place
            DeviceVector<double> xd{HostVector<double>(bv)};
            M->MInvSolve(xd);  // device vector, in place
            HostVector<double> back(n);
            xd.copyToHost(back);
            d_host_dev = maxdiff(xh, back);
            // offset form on a longer host vector
            HostVector<double> longer(n + 7);
            for (uint32_t i = 0; i < n; ++i) longer[i + 5] = bv[i];
            M->MInvSolve(longer, 5);
            d_offset = maxdiff(xh, longer, 5);
            // in -> out form leaves the input alone
            HostVector<double> in(bv), out;
            M->MInvSolve(in, out);
            d_inout = std::fmax(maxdiff(xh, out), maxdiff(in, HostVector<double>(bv)));
            // MSolveLower, MSolveDiagonal, MSolveUpper compose to MInvSolve
            auto *T = dynamic_cast<DeviceTriangularPrecondition *>(M.get());
            if (T) {
                HostVector<double> p(bv);
                T->MSolveLower(p);
                T->MSolveDiagonal(p);
                T->MSolveUpper(p);
                double s = 0.0;
                for (uint32_t i = 0; i < n; ++i) s = std::fmax(s, std::fabs(xh[i]));
                d_parts = maxdiff(xh, p) / (s > 0 ? s : 1.0);
            }
        }

        // host drop-in solve: BaseSolver<double>::solve(A, b, x)
        std::unique_ptr<DeviceSolver> S;
        if (solver == "cg") S.reset(new DeviceCG(n, M, FlexibleCG, tol, 1000));
        else {
            auto *g = new DeviceGMRES(n, M, tol, FlexibleGMRES);
            g->setRestart(true, 30);
            S.reset(g);
        }
        BaseSolver<double> &base = *S;
        HostVector<double> b(bv), x(n);
        base.solve(A, b, x);
        const int it_host = S->getNumIter();
        const double err_host = S->getError();
        const bool conv_host = S->getConvergence();

        // device-resident solve on the same objects
        DeviceVector<double> db(b), dx(n);
        S->solve(*dA, db, dx);
        HostVector<double> xdev(n);
        dx.copyToHost(xdev);

        std::ofstream o(argv[2], std::ios::binary);
        o.write(reinterpret_cast<const char *>(x.getRawValPtr()), 8ll * n);
        o.write(reinterpret_cast<const char *>(xdev.getRawValPtr()), 8ll * n);

        // an invalid call must throw, not exit
        bool threw = false;
        try {
            HostVector<double> shortb(n > 1 ? n - 1 : 0), xx(n);
            base.solve(A, shortb, xx);
        } catch (const Error &e) {
            threw = e.status() == PP_ERR_INVALID;
        }
        std::printf(
            "{\"iters\": %d, \"error\": %.17g, \"converged\": %d, \"iters_device\": %d, \"converged_device\": %d, "
            "\"ready_before\": %d, \"ready_after\": %d, \"fill\": %.17g, \"d_host_dev\": %.17g, \"d_offset\": %.17g, "
            "\"d_inout\": %.17g, \"d_parts\": %.17g, \"type\": %d, \"invalid_threw\": %d}\n",
            it_host, err_host, conv_host ? 1 : 0, S->getNumIter(), S->getConvergence() ? 1 : 0, ready_before,
            ready_after, fill, d_host_dev, d_offset, d_inout, d_parts, (int)S->getPrecondtionType(), threw ? 1 : 0);
        return 0;
    } catch (const Error &e) {
        std::printf("{\"exception\": \"%s\", \"status\": %d}\n", e.what(), e.status());
        return 1;
    } catch (const std::exception &e) {
        std::printf("{\"exception\": \"%s\", \"status\": -1}\n", e.what());
        return 1;
    }
}
