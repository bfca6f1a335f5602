// A compiled expression: the postfix program of an Expr (expr.h), evaluated
// by the same inline interpreter on the host and on the device.  The mask
// formula of a masked run is evaluated at every plane of every column (the
// interface weights, UtilitiesLib/VectorField.cpp:64-113); on the device one
// thread takes one slab of one column.
//
// Arithmetic is IEEE double with one rounding per operation (no contraction
// across operations: every result goes through the evaluation stack), so
// + - * / and x^2 (evaluated as x*x, as muParser's byte code does) give the
// same bits on both sides; the elementary functions come from the host's and
// the device's math libraries and may differ in the last place.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace nmk {

enum ExprCode : int32_t {
    EX_CONST = 0, EX_VAR, EX_NEG, EX_SQR, EX_FUN1,
    EX_ADD, EX_SUB, EX_MUL, EX_DIV, EX_POW, EX_ATAN2,
    EX_LT, EX_GT, EX_LE, EX_GE, EX_EQ, EX_NE, EX_AND, EX_OR,
    EX_TERN, EX_SUM, EX_AVG, EX_MIN, EX_MAX
};
enum ExprFun : int32_t {
    EF_SIN = 0, EF_COS, EF_TAN, EF_ASIN, EF_ACOS, EF_ATAN, EF_SINH, EF_COSH, EF_TANH, EF_ASINH, EF_ACOSH,
    EF_ATANH, EF_LOG2, EF_LOG10, EF_LOG, EF_EXP, EF_SQRT, EF_ABS, EF_SIGN, EF_RINT
};

struct ExprOp {
    int32_t code, arg;  // arg: variable index, function, or argument count
    double value;
};
constexpr int kExprMaxOps = 96, kExprMaxStack = 24;
struct ExprProgram {
    int32_t n = 0, depth = 0;
    ExprOp ops[kExprMaxOps];
};

__host__ __device__ inline double expr_fun1(int f, double a) {
    switch (f) {
        case EF_SIN: return sin(a);
        case EF_COS: return cos(a);
        case EF_TAN: return tan(a);
        case EF_ASIN: return asin(a);
        case EF_ACOS: return acos(a);
        case EF_ATAN: return atan(a);
        case EF_SINH: return sinh(a);
        case EF_COSH: return cosh(a);
        case EF_TANH: return tanh(a);
        case EF_ASINH: return asinh(a);
        case EF_ACOSH: return acosh(a);
        case EF_ATANH: return atanh(a);
        case EF_LOG2: return log2(a);
        case EF_LOG10: return log10(a);
        case EF_LOG: return log(a);
        case EF_EXP: return exp(a);
        case EF_SQRT: return sqrt(a);
        case EF_ABS: return fabs(a);
        case EF_SIGN: return a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0);
        case EF_RINT: return rint(a);
    }
    return 0;
}

// evaluation stack of one evaluation: a local array, or (on the device) a
// thread's column of a block-wide LDS array (slot k of thread t at k*stride+t,
// so that a wave's accesses fall on consecutive banks)
struct ExprArrayStack {
    double *s;
    __host__ __device__ double &operator[](int k) const { return s[k]; }
};
struct ExprStridedStack {
    double *s;
    int stride;
    __host__ __device__ double &operator[](int k) const { return s[k * stride]; }
};

// value of the program for the variables (x0, x1, x2) (in the order the
// expression was compiled with; unused ones are ignored); st holds at least
// p.depth slots
template <class Stack>
__host__ __device__ inline double expr_eval(const ExprProgram &p, double x0, double x1, double x2, Stack st) {
    int sp = 0;
    for (int i = 0; i < p.n; i++) {
        const ExprOp op = p.ops[i];
        switch (op.code) {
            case EX_CONST: st[sp++] = op.value; break;
            case EX_VAR: st[sp++] = op.arg == 0 ? x0 : op.arg == 1 ? x1 : x2; break;
            case EX_NEG: st[sp - 1] = -st[sp - 1]; break;
            case EX_SQR: {
                const double a = st[sp - 1];
                st[sp - 1] = a * a;
                break;
            }
            case EX_FUN1: st[sp - 1] = expr_fun1(op.arg, st[sp - 1]); break;
            case EX_TERN: {
                const double c = st[sp - 3], a = st[sp - 2], b = st[sp - 1];
                sp -= 2;
                st[sp - 1] = c != 0 ? a : b;
                break;
            }
            case EX_SUM:
            case EX_AVG: {
                const int n = op.arg;
                double s = 0;
                for (int k = sp - n; k < sp; k++) s += st[k];
                sp -= n;
                st[sp++] = op.code == EX_SUM ? s : s / double(n);
                break;
            }
            case EX_MIN:
            case EX_MAX: {
                const int n = op.arg;
                double m = st[sp - n];
                for (int k = sp - n; k < sp; k++) {
                    const double x = st[k];
                    m = op.code == EX_MIN ? (x < m ? x : m) : (x > m ? x : m);
                }
                sp -= n;
                st[sp++] = m;
                break;
            }
            default: {
                const double b = st[--sp], a = st[sp - 1];
                double r = 0;
                switch (op.code) {
                    case EX_ADD: r = a + b; break;
                    case EX_SUB: r = a - b; break;
                    case EX_MUL: r = a * b; break;
                    case EX_DIV: r = a / b; break;
                    case EX_POW: r = pow(a, b); break;
                    case EX_ATAN2: r = atan2(a, b); break;
                    case EX_LT: r = a < b; break;
                    case EX_GT: r = a > b; break;
                    case EX_LE: r = a <= b; break;
                    case EX_GE: r = a >= b; break;
                    case EX_EQ: r = a == b; break;
                    case EX_NE: r = a != b; break;
                    case EX_AND: r = (a != 0 && b != 0) ? 1 : 0; break;
                    case EX_OR: r = (a != 0 || b != 0) ? 1 : 0; break;
                }
                st[sp - 1] = r;
            }
        }
    }
    return sp > 0 ? st[0] : 0;
}

// host convenience: the three variables from v[0..2], a local stack
inline double expr_run(const ExprProgram &p, const double *v) {
    double s[kExprMaxStack];
    return expr_eval(p, v[0], v[1], v[2], ExprArrayStack{s});
}

}  // namespace nmk
