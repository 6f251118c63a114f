// precn.hpp -- preclib's precn_t interface (prec.hpp:49-205) over libpgpu.so.
//
// A drop-in for the product path: the type keeps preclib's layout and
// malloc-owned lifecycle (prec.hpp:49-84, src/init_basic.cpp:47-135), and
// operator*, mul_into, precn_sqr, mul_ntt and mul_fft keep their names and
// results, with every transform product on the GPU.  So do the callers on the
// path: Newton division (div_mulinv, mod_mulinv, precn_reciprocal_newton,
// src/div_newton.cpp), precn_sqrt (src/sqrt.cpp) and base-10 conversion
// (src/base_convert.cpp).  The reference never fails; here a library error
// (no device, budget exhausted) throws precn_gpu_error.  There is no CPU
// product path behind it.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>

struct precn_gpu_error : std::runtime_error {
    int code;
    precn_gpu_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Opens the device (pg_init); called implicitly with device 0 on first use.
void precn_gpu_init(int device = 0, size_t hbm_budget = 0);

struct precn_t {   // unsigned arbitrary precision number, base 2^64 (prec.hpp:49-84)
    size_t asiz;   // allocated limbs
    size_t rsiz;   // used limbs (trimmed: a[rsiz - 1] != 0)
    uint64_t *a;   // malloc-owned limbs, little endian
    precn_t();
    template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
    precn_t(T val) {
        a = (uint64_t *)malloc(sizeof(uint64_t));
        asiz = 1;
        rsiz = 0;
        *a = 0;
        if (std::is_signed<T>::value && val < 0) std::abort();
        if (val) {
            a[0] = (uint64_t)val;
            rsiz = 1;
        }
    }
    precn_t(const precn_t &o);
    precn_t(precn_t &&o);
    template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
    explicit operator T() const {
        return rsiz == 0 ? (T)0 : (T)a[0];
    }
    explicit operator std::string() const;   // decimal, "0" for zero
    precn_t &operator=(const precn_t &o);
    precn_t &operator=(precn_t &&o);
    ~precn_t();
};

// ---- products (src/mul_basic.cpp, mul_ntt.cpp, mul_fft.cpp) ----------------------
precn_t operator*(const precn_t &a, const precn_t &b);
void mul_into(precn_t &r, const precn_t &a, const precn_t &b);
precn_t precn_sqr(const precn_t &a);
precn_t mul_ntt(const precn_t &a, const precn_t &b);
precn_t mul_fft(const precn_t &a, const precn_t &b);
// mul_fft plus its largest rounding distance (the COUNT_FFTS statistic).
precn_t mul_fft_round(const precn_t &a, const precn_t &b, double *max_round);
precn_t mul_u32(const precn_t &a, uint32_t b);
precn_t mul_u64(const precn_t &a, uint64_t b);

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator*(const precn_t &a, T b) {
    if (std::is_signed<T>::value && b < 0) std::abort();
    return mul_u64(a, (uint64_t)b);
}

// ---- division, square root, conversion -------------------------------------------
precn_t div_mulinv(const precn_t &a, const precn_t &b);
precn_t mod_mulinv(const precn_t &a, const precn_t &b);
precn_t precn_reciprocal_newton(const precn_t &b, size_t bits);
precn_t operator/(const precn_t &a, const precn_t &b);
precn_t operator%(const precn_t &a, const precn_t &b);
precn_t precn_sqrt(const precn_t &a);
// base 10 or a power of two; out gets out_siz digits, least significant first
void precn_base_convert(const precn_t &a, uint32_t base, uint32_t *out, size_t &out_siz);

// ---- linear steps (src/add.cpp, sub.cpp, compare_shift.cpp) ------------------------
precn_t operator+(const precn_t &a, const precn_t &b);
precn_t operator-(const precn_t &a, const precn_t &b);   // requires a >= b
precn_t add_u64(const precn_t &a, uint64_t b);
precn_t operator<<(const precn_t &a, size_t bits);
precn_t operator>>(const precn_t &a, size_t bits);
int precn_cmp(const precn_t &a, const precn_t &b);
inline bool operator==(const precn_t &a, const precn_t &b) { return precn_cmp(a, b) == 0; }
inline bool operator!=(const precn_t &a, const precn_t &b) { return precn_cmp(a, b) != 0; }
inline bool operator<(const precn_t &a, const precn_t &b) { return precn_cmp(a, b) < 0; }
inline bool operator>(const precn_t &a, const precn_t &b) { return precn_cmp(a, b) > 0; }
inline bool operator<=(const precn_t &a, const precn_t &b) { return precn_cmp(a, b) <= 0; }
inline bool operator>=(const precn_t &a, const precn_t &b) { return precn_cmp(a, b) >= 0; }
