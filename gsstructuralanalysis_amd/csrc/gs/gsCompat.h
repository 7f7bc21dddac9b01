// gsCompat.h -- the subset of gismo's linear-algebra / option / solver types that the
// structural-analysis solvers use, re-expressed over libksh.so device objects.
//
//   gsVector<T>        host vector (n <= ~1e5: dots/axpys on the host are cheap; K stays on device)
//   gsSparseMatrix<T>  handle to a device band matrix (ksh_matrix) of one problem: the solvers
//                      only apply compute/solve, SpMV, A - s B, B -= A, rows/cols to it
//                      (src/gsALMSolvers/gsALMBase.hpp:639,663,720; gsEigenProblemBase.hpp:32,105;
//                      gsBucklingSolver.h:104,144)
//   gsSparseSolver<T>  ::get(name) -> SimplicialLDLT (band LDL^T on the GPU; vectorD())
//                      (src/gsALMSolvers/gsALMBase.hpp:73-78,186-211,559-562)
//   gsOptionList       add/get/set Int/Real/Switch/String, askString, update
//
// Matrices are recycled through a per-problem pool so the solvers' value semantics
// (`gsSparseMatrix<T> m; jacobian(u, m); return m;`) cost no device allocation per call.
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/ksh.h"

namespace gismo {

typedef int index_t;
typedef double real_t;

#define gsInfo std::cout
#define gsWarn std::cerr << "Warning: "

namespace math {
template <class T> inline T pow(T a, T b) { return std::pow(a, b); }
template <class T> inline T sqrt(T a) { return std::sqrt(a); }
template <class T> inline T abs(T a) { return std::fabs(a); }
template <class T> inline T log10(T a) { return std::log10(a); }
}  // namespace math

// ------------------------------------------------------------------------------------------
template <class T>
class gsVector {
public:
    gsVector() = default;
    explicit gsVector(index_t n) : m_v((size_t)n, T(0)) {}
    gsVector(std::vector<T> v) : m_v(std::move(v)) {}

    static gsVector Zero(index_t n) { return gsVector(n); }
    static gsVector Ones(index_t n)
    {
        gsVector v(n);
        for (auto &x : v.m_v) x = T(1);
        return v;
    }

    index_t size() const { return (index_t)m_v.size(); }
    index_t rows() const { return (index_t)m_v.size(); }
    index_t cols() const { return m_v.empty() ? 0 : 1; }
    void resize(index_t n) { m_v.resize((size_t)n); }
    void setZero() { std::fill(m_v.begin(), m_v.end(), T(0)); }
    void setZero(index_t n) { m_v.assign((size_t)n, T(0)); }
    void setOnes() { std::fill(m_v.begin(), m_v.end(), T(1)); }
    T *data() { return m_v.data(); }
    const T *data() const { return m_v.data(); }
    T &operator[](index_t i) { return m_v[(size_t)i]; }
    const T &operator[](index_t i) const { return m_v[(size_t)i]; }
    T &operator()(index_t i) { return m_v[(size_t)i]; }
    const T &operator()(index_t i) const { return m_v[(size_t)i]; }
    T at(index_t i) const { return m_v.at((size_t)i); }

    T dot(const gsVector &o) const
    {
        check(o);
        T s = 0;
        for (size_t i = 0; i < m_v.size(); ++i) s += m_v[i] * o.m_v[i];
        return s;
    }
    T squaredNorm() const { return dot(*this); }
    T norm() const { return std::sqrt(squaredNorm()); }
    void normalize()
    {
        T n = norm();
        if (n > 0)
            for (auto &x : m_v) x /= n;
    }
    T minCoeff() const
    {
        T m = m_v.empty() ? T(0) : m_v[0];
        for (auto x : m_v) m = x < m ? x : m;
        return m;
    }

    gsVector &operator+=(const gsVector &o)
    {
        check(o);
        for (size_t i = 0; i < m_v.size(); ++i) m_v[i] += o.m_v[i];
        return *this;
    }
    gsVector &operator-=(const gsVector &o)
    {
        check(o);
        for (size_t i = 0; i < m_v.size(); ++i) m_v[i] -= o.m_v[i];
        return *this;
    }
    gsVector &operator*=(T a)
    {
        for (auto &x : m_v) x *= a;
        return *this;
    }
    gsVector &operator/=(T a)
    {
        for (auto &x : m_v) x /= a;
        return *this;
    }
    friend gsVector operator+(gsVector a, const gsVector &b) { return a += b; }
    friend gsVector operator-(gsVector a, const gsVector &b) { return a -= b; }
    friend gsVector operator*(T s, gsVector a) { return a *= s; }
    friend gsVector operator*(gsVector a, T s) { return a *= s; }
    friend gsVector operator/(gsVector a, T s) { return a /= s; }
    gsVector operator-() const
    {
        gsVector r(*this);
        for (auto &x : r.m_v) x = -x;
        return r;
    }
    const std::vector<T> &std() const { return m_v; }

private:
    void check(const gsVector &o) const
    {
        if (o.m_v.size() != m_v.size()) throw std::length_error("gsVector size mismatch");
    }
    std::vector<T> m_v;
};

// ------------------------------------------------------------------------------------------
// Executor of the device operations behind the operator surface.  The default runs each call
// on libksh.so directly; a batch engine supplies one that defers the call into a batched launch
// over many problems (the calling solver blocks until its result is ready).
struct gsExec {
    virtual ~gsExec() = default;
    virtual int tangent(ksh_problem *pb, const double *u, ksh_matrix *K) { return ksh_assemble_tangent(pb, u, K); }
    virtual int residual(ksh_problem *pb, const double *u, double lam, int conv, double *R)
    {
        return ksh_assemble_residual(pb, u, lam, conv, R);
    }
    virtual int factor(ksh_matrix *K) { return ksh_factor(K); }
    virtual int solve(ksh_matrix *K, const double *b, double *x, int max_refine, double tol)
    {
        return ksh_solve(K, b, x, max_refine, tol);
    }
    virtual int inertia(ksh_matrix *K, int64_t *neg, double *dmin) { return ksh_inertia(K, neg, dmin); }
    virtual int matvec(const ksh_matrix *K, const double *x, double *y) { return ksh_matvec(K, x, y); }
    virtual int vector_d(ksh_matrix *K, double *d) { return ksh_vector_d(K, d); }
    // learned predictor network, one sample (x: network inputs, y: outputs)
    virtual int predict(ksh_mlp *m, const double *x, double *y) { return ksh_mlp_predict(m, x, 1, y); }
    static gsExec *direct()
    {
        static gsExec e;
        return &e;
    }
};

// Device band matrix handle + pool
class gsMatrixPool;

struct gsMatrixStore {
    ksh_matrix *K = nullptr;
    ksh_problem *pb = nullptr;
    int prec = KSH_FP64;
    gsExec *exec = gsExec::direct();
    ~gsMatrixStore()
    {
        if (K) ksh_matrix_destroy(K);
    }
};

template <class T>
class gsSparseMatrix {
public:
    gsSparseMatrix() = default;
    explicit gsSparseMatrix(std::shared_ptr<gsMatrixStore> s) : m_s(std::move(s)) {}

    bool valid() const { return m_s && m_s->K; }
    ksh_matrix *handle() const { return m_s ? m_s->K : nullptr; }
    ksh_problem *problem() const { return m_s ? m_s->pb : nullptr; }
    gsExec *exec() const { return m_s ? m_s->exec : gsExec::direct(); }
    std::shared_ptr<gsMatrixStore> store() const { return m_s; }
    index_t rows() const
    {
        int64_t n = 0;
        if (valid()) ksh_matrix_dims(m_s->K, &n, nullptr, nullptr);
        return (index_t)n;
    }
    index_t cols() const { return rows(); }

    // y = K x on the device
    gsVector<T> operator*(const gsVector<T> &x) const
    {
        gsVector<T> y(rows());
        if (exec()->matvec(handle(), x.data(), y.data()) != KSH_SUCCESS) throw std::runtime_error(ksh_last_error());
        return y;
    }

    // deep copy into a new device matrix of the same problem and precision
    gsSparseMatrix clone() const
    {
        auto s = std::make_shared<gsMatrixStore>();
        s->pb = m_s->pb;
        s->prec = m_s->prec;
        s->exec = m_s->exec;
        if (ksh_matrix_create(s->pb, s->prec, &s->K) != KSH_SUCCESS || ksh_matrix_copy(s->K, m_s->K) != KSH_SUCCESS)
            throw std::runtime_error(ksh_last_error());
        return gsSparseMatrix(s);
    }
    // this = alpha this + beta o, in place (A - sigma B and B -= A of
    // src/gsEigenSolvers/gsEigenProblemBase.hpp:32 and gsBucklingSolver.h:104,144)
    gsSparseMatrix &axpby(T alpha, T beta, const gsSparseMatrix &o)
    {
        if (ksh_matrix_axpby(handle(), alpha, beta, o.handle()) != KSH_SUCCESS) throw std::runtime_error(ksh_last_error());
        return *this;
    }
    gsSparseMatrix &operator-=(const gsSparseMatrix &o) { return axpby(T(1), T(-1), o); }
    gsSparseMatrix &operator+=(const gsSparseMatrix &o) { return axpby(T(1), T(1), o); }
    friend gsSparseMatrix operator-(const gsSparseMatrix &a, const gsSparseMatrix &b)
    {
        gsSparseMatrix r = a.clone();
        r -= b;
        return r;
    }

private:
    std::shared_ptr<gsMatrixStore> m_s;
};

// Pool of device matrices of one problem and precision; a store is reused once no
// gsSparseMatrix refers to it any more.
class gsMatrixPool {
public:
    gsMatrixPool(ksh_problem *pb, int prec, gsExec *exec = gsExec::direct()) : m_pb(pb), m_prec(prec), m_exec(exec) {}
    template <class T> gsSparseMatrix<T> acquire()
    {
        for (auto &s : m_stores)
            if (s.use_count() == 1) return gsSparseMatrix<T>(s);
        auto s = std::make_shared<gsMatrixStore>();
        s->pb = m_pb;
        s->prec = m_prec;
        s->exec = m_exec;
        if (ksh_matrix_create(m_pb, m_prec, &s->K) != KSH_SUCCESS) throw std::runtime_error(ksh_last_error());
        m_stores.push_back(s);
        return gsSparseMatrix<T>(s);
    }
    size_t size() const { return m_stores.size(); }
    ksh_problem *problem() const { return m_pb; }
    int precision() const { return m_prec; }
    gsExec *exec() const { return m_exec; }

private:
    ksh_problem *m_pb;
    int m_prec;
    gsExec *m_exec;
    std::vector<std::shared_ptr<gsMatrixStore>> m_stores;
};

// ------------------------------------------------------------------------------------------
namespace gsEigen {
enum ComputationInfo { Success = 0, NumericalIssue = 1, NoConvergence = 2, InvalidInput = 3 };
}

// Settings of the device solves used by the LDL^T adaptor (iterative refinement)
struct gsSolveSettings {
    int max_refine = 0;
    double refine_tol = 1e-14;
};

template <class T>
class gsSparseSolver {
public:
    typedef std::unique_ptr<gsSparseSolver> uPtr;
    class SimplicialLDLT;
    class CGDiagonal;
    virtual ~gsSparseSolver() = default;

    static uPtr get(const std::string &name);

    virtual void compute(const gsSparseMatrix<T> &K)
    {
        m_K = K;
        int rc = K.exec()->factor(K.handle());
        m_info = rc == KSH_SUCCESS ? gsEigen::Success : gsEigen::NumericalIssue;
    }
    virtual gsVector<T> solve(const gsVector<T> &b) const
    {
        gsVector<T> x(b.rows());
        int rc = m_K.exec()->solve(m_K.handle(), b.data(), x.data(), settings().max_refine, settings().refine_tol);
        if (rc != KSH_SUCCESS) throw std::runtime_error(ksh_last_error());
        return x;
    }
    gsEigen::ComputationInfo info() const { return m_info; }
    const gsSparseMatrix<T> &matrix() const { return m_K; }

    static gsSolveSettings &settings()
    {
        static thread_local gsSolveSettings s;
        return s;
    }

protected:
    gsSparseMatrix<T> m_K;
    gsEigen::ComputationInfo m_info = gsEigen::InvalidInput;
};

template <class T>
class gsSparseSolver<T>::SimplicialLDLT : public gsSparseSolver<T> {
public:
    // diagonal D of the LDL^T factor; its negative count is the inertia of K
    gsVector<T> vectorD() const
    {
        gsVector<T> d(this->m_K.rows());
        if (this->m_K.exec()->vector_d(this->m_K.handle(), d.data()) != KSH_SUCCESS) throw std::runtime_error(ksh_last_error());
        return d;
    }
    // inertia with the FP32 near-zero-pivot guard (FP64 refactorization)
    int64_t negatives(T *dmin) const
    {
        int64_t neg = 0;
        double dm = 0;
        if (this->m_K.exec()->inertia(this->m_K.handle(), &neg, &dm) != KSH_SUCCESS) throw std::runtime_error(ksh_last_error());
        if (dmin) *dmin = (T)dm;
        return neg;
    }
};

// "CGDiagonal" is served by the same exact band solve; it is a distinct type so that the
// determinant stability method stays restricted to SimplicialLDLT, as in the reference.
template <class T>
class gsSparseSolver<T>::CGDiagonal : public gsSparseSolver<T> {};

template <class T>
typename gsSparseSolver<T>::uPtr gsSparseSolver<T>::get(const std::string &name)
{
    if (name == "SimplicialLDLT") return uPtr(new SimplicialLDLT());
    if (name == "CGDiagonal") return uPtr(new CGDiagonal());
    throw std::invalid_argument("gsSparseSolver: unsupported solver '" + name + "'");
}

// ------------------------------------------------------------------------------------------
class gsOptionList {
public:
    enum { addIfUnknown = 1 };
    void addInt(const std::string &k, const std::string &, int v) { m_int[k] = v; }
    void addReal(const std::string &k, const std::string &, double v) { m_real[k] = v; }
    void addSwitch(const std::string &k, const std::string &, bool v) { m_bool[k] = v; }
    void addString(const std::string &k, const std::string &, const std::string &v) { m_str[k] = v; }
    void setInt(const std::string &k, int v) { m_int.at(k) = v; }
    void setReal(const std::string &k, double v) { m_real.at(k) = v; }
    void setSwitch(const std::string &k, bool v) { m_bool.at(k) = v; }
    void setString(const std::string &k, const std::string &v) { m_str.at(k) = v; }
    int getInt(const std::string &k) const { return m_int.at(k); }
    double getReal(const std::string &k) const { return m_real.at(k); }
    bool getSwitch(const std::string &k) const { return m_bool.at(k); }
    std::string getString(const std::string &k) const { return m_str.at(k); }
    std::string askString(const std::string &k, const std::string &def) const
    {
        auto it = m_str.find(k);
        return it == m_str.end() ? def : it->second;
    }
    bool hasInt(const std::string &k) const { return m_int.count(k) != 0; }
    bool hasReal(const std::string &k) const { return m_real.count(k) != 0; }
    void update(const gsOptionList &o, int)
    {
        for (auto &kv : o.m_int) m_int[kv.first] = kv.second;
        for (auto &kv : o.m_real) m_real[kv.first] = kv.second;
        for (auto &kv : o.m_bool) m_bool[kv.first] = kv.second;
        for (auto &kv : o.m_str) m_str[kv.first] = kv.second;
    }

private:
    std::map<std::string, int> m_int;
    std::map<std::string, double> m_real;
    std::map<std::string, bool> m_bool;
    std::map<std::string, std::string> m_str;
};

}  // namespace gismo
