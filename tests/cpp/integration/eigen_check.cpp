// The Eigen conversions of include/ipc_amd.hpp against the Eigen stand-in
// (no device needed): dense round trip with the storage orders swapped,
// gradient vector, Hessian CSR -> Eigen's compressed format.
#include <Eigen/Core>
#include <ipc_amd.hpp>

#include <cstdio>

int main()
{
    int failures = 0;
    Eigen::MatrixXd V(4, 3);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 3; j++) V(i, j) = 10.0 * i + j + 0.5;
    const ipc::MatrixXd M = ipc::from_eigen(V);
    const Eigen::MatrixXd W = ipc::to_eigen(M);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 3; j++) failures += M(i, j) != V(i, j) || W(i, j) != V(i, j);
    Eigen::MatrixXi E(2, 2);
    E(0, 0) = 0; E(0, 1) = 1; E(1, 0) = 1; E(1, 1) = 2;
    const ipc::MatrixXi Ei = ipc::from_eigen(E);
    failures += Ei(1, 0) != 1 || Ei(1, 1) != 2;
    const Eigen::VectorXd g = ipc::to_eigen(std::vector<double> { 1.0, -2.0, 3.0 });
    failures += g.size() != 3 || g.data()[1] != -2.0;
    ipc::SparseMatrix H;
    H.rows = H.cols = 3;
    H.outer = { 0, 2, 3, 4 };
    H.inner = { 0, 2, 1, 0 };
    H.values = { 4.0, 1.0, 5.0, 1.0 };
    Eigen::SparseMatrix<double> S = ipc::to_eigen(H);
    failures += S.nonZeros() != 4 || S.outerIndexPtr()[3] != 4 || S.innerIndexPtr()[1] != 2
        || S.valuePtr()[2] != 5.0;
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
