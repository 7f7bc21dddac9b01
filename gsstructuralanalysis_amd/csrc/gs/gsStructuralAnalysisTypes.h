// gsStructuralAnalysisTypes.h -- the operator surface of the structural-analysis solvers:
// gsStatus and the std::function typedefs of gsStructuralAnalysisOps
// (src/gsStructuralAnalysisTools/gsStructuralAnalysisTypes.h:22-30,58-93), over the compat
// vector / device-matrix types.  A Jacobian_t fills a device-resident band matrix handle;
// residuals are host vectors.
#pragma once

#include <functional>

#include "gsCompat.h"

namespace gismo {

enum struct gsStatus { Success, NotConverged, AssemblyError, SolverError, NotStarted, OtherError };

inline gsStatus status_from_code(int code)
{
    switch (code) {
    case 1: return gsStatus::NotConverged;
    case 2: return gsStatus::AssemblyError;
    case 3: return gsStatus::SolverError;
    default: return gsStatus::OtherError;
    }
}

template <class T>
struct gsStructuralAnalysisOps {
    typedef std::function<bool(gsVector<T> const &, T &)> Energy_t;
    typedef std::function<bool(gsVector<T> &)> Force_t;
    typedef std::function<bool(const T, gsVector<T> &)> TForce_t;
    /// residual (static convention: F_ext - F_int)
    typedef std::function<bool(gsVector<T> const &, gsVector<T> &)> Residual_t;
    /// arc-length residual (F_int - lambda F_ext)
    typedef std::function<bool(gsVector<T> const &, const T, gsVector<T> &)> ALResidual_t;
    typedef std::function<bool(gsVector<T> const &, const T, gsVector<T> &)> TResidual_t;
    typedef std::function<bool(gsSparseMatrix<T> &)> Mass_t;
    typedef std::function<bool(const T, gsSparseMatrix<T> &)> TMass_t;
    typedef std::function<bool(gsVector<T> const &, gsSparseMatrix<T> &)> Damping_t;
    typedef std::function<bool(gsVector<T> const &, const T, gsSparseMatrix<T> &)> TDamping_t;
    typedef std::function<bool(gsSparseMatrix<T> &)> Stiffness_t;
    typedef std::function<bool(gsVector<T> const &, gsSparseMatrix<T> &)> Jacobian_t;
    typedef std::function<bool(gsVector<T> const &, const T, gsSparseMatrix<T> &)> TJacobian_t;
    typedef std::function<bool(gsVector<T> const &, gsVector<T> const &, gsSparseMatrix<T> &)> dJacobian_t;
};

}  // namespace gismo
