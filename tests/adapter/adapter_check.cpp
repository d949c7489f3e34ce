// Compile check of include/ug_gpu/ug_bridge_adapter.hpp against the
// reference's own interface headers (ugbase/ on the include path): the GPU1
// classes derive from ILinearIterator, IPreconditioner, ILinearOperator and
// MatrixOperator, and the reference's generic code instantiates with them.
#include "ug_gpu/ug_bridge_adapter.hpp"

#include "lib_algebra/algebra_common/sparsematrix_util.h"
#include "lib_algebra/common/operations_mat/operations_mat.h"

namespace ug {
// every member of the reference's class templates, instantiated for GPU1
// (IPreconditioner<GPUAlgebra> is instantiated through GPUJacobi below: its
// non-virtual defect_operator() does not compile for any algebra, CPUAlgebra
// included, so the whole class template cannot be instantiated explicitly)
template class ILinearIterator<GPUVector>;
template class ILinearOperator<GPUVector>;
template class MatrixOperator<GPUSparseMatrix, GPUVector>;
template class ConstantDamping<GPUVector>;
}  // namespace ug

using namespace ug;

// what a bridge registration and a script's set-up do with the classes
bool exercise(void* disc) {
    SmartPtr<GPUMatrixOperator> A = make_sp(new GPUAssembledLinearOperator(disc));
    A->init();
    GPUVector x(A->num_rows()), b(A->num_rows());
    b.set(1.0);
    x = 0.0;

    SmartPtr<GPUJacobi> jac = make_sp(new GPUJacobi(0.66));
    SmartPtr<IPreconditioner<GPUAlgebra> > pre = jac;
    SmartPtr<ILinearIterator<GPUVector> > it = pre;
    SmartPtr<ILinearOperator<GPUVector> > L = A;
    if (!it->init(L, x)) return false;
    if (!it->apply(x, b)) return false;
    if (!it->apply_update_defect(x, b)) return false;
    SmartPtr<ILinearIterator<GPUVector> > it2 = it->clone();

    SmartPtr<GPUGeometricMultiGrid> gmg = make_sp(new GPUGeometricMultiGrid(disc));
    gmg->set_base_level(0);
    gmg->set_cycle_type("F");
    gmg->set_num_presmooth(2);
    gmg->set_num_postsmooth(2);
    gmg->set_smoother(jac);
    gmg->set_damp(1.0);
    SmartPtr<ILinearIterator<GPUVector> > mg = gmg;
    if (!mg->init(L)) return false;
    if (!mg->apply_update_defect(x, b)) return false;

    // the matrix surface the reference's generic code reaches
    GPUSparseMatrix& M = *A;
    SetDirichletRow(M, 0);
    MatMultAdd(b, 1.0, b, -1.0, M, x);
    MatMultTransposedAdd(b, 1.0, b, 2.0, M, x);
    number s = 0.0;
    for (GPUSparseMatrix::const_row_iterator c = M.begin_row(1); c != M.end_row(1); ++c) s += c.value() * c.index();
    GPUSparseMatrix T;
    T.set_as_transpose_of(M);
    T.resize_and_clear(3, 3);
    T(0, 0) += 2.0;
    T(1, 2) = -1.0;
    T.defragment();
    VecScaleAdd(x, 1.0, x, s, b);
    VecScaleAdd(x, 1.0, x, 0.5, b, 0.25, b);
    x[0] = 3.0;
    number first = x[0];
    return VecProd(x, b) + x.norm() + first > 0.0;
}

#include <type_traits>
static_assert(std::is_base_of<IPreconditioner<GPUAlgebra>, GPUJacobi>::value, "Jacobi is an IPreconditioner");
static_assert(std::is_base_of<ILinearIterator<GPUVector>, GPUJacobi>::value, "Jacobi is an ILinearIterator");
static_assert(std::is_base_of<ILinearIterator<GPUVector>, GPUGeometricMultiGrid>::value, "GMG is an ILinearIterator");
static_assert(std::is_base_of<ILinearOperator<GPUVector>, GPUMatrixOperator>::value, "a matrix operator");
static_assert(std::is_base_of<GPUSparseMatrix, GPUAssembledLinearOperator>::value, "an assembled operator is a matrix");
static_assert(!std::is_abstract<GPUJacobi>::value && !std::is_abstract<GPUGeometricMultiGrid>::value &&
                  !std::is_abstract<GPUAssembledLinearOperator>::value,
              "every pure virtual of the reference interfaces is implemented");
static_assert(GPUAlgebra::blockSize == 1, "GPU1");
