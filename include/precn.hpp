// precn.hpp -- preclib's public interface (prec.hpp) over libpgpu.so.
//
// A drop-in for prec.hpp: the same types (precn_t with preclib's layout and
// malloc-owned lifecycle, prec.hpp:49-84 / src/init_basic.cpp:47-135; the
// signed precz_t, prec.hpp:207-289), the same inline limb primitives
// (prec.hpp:14-47) and every function prec.hpp declares, with the same
// names, signatures and results.  Every product (operator*, mul_into,
// precn_sqr, mul_ntt, mul_fft and the reference's alternative multipliers)
// runs on the GPU through pg_mul; divisions by a divisor of two or more limbs
// run Newton's method on the GPU (pg_div); decimal conversion both ways, the
// square root and binary splitting run there too.  Linear-time steps
// (addition, subtraction, shifts, comparisons, products and quotients by one
// machine word) stay on the host, as in the reference.
//
// The reference never fails; here a library error (no gfx950 device, device
// memory exhausted) throws precn_gpu_error.  There is no CPU product path.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

#if defined(__SIZEOF_INT128__)
typedef unsigned __int128 uint128_t;
#endif

// ---- limb primitives (prec.hpp:14-47) ---------------------------------------------
// out = a + b + carry; returns the carry out.
static inline uint64_t precn_add_carry(uint64_t a, uint64_t b, uint64_t carry, uint64_t &out) {
    const unsigned __int128 s = (unsigned __int128)a + b + carry;
    out = (uint64_t)s;
    return (uint64_t)(s >> 64);
}

// out = a - b - borrow (mod 2^64); returns 1 when b + borrow exceeds a.
static inline uint64_t precn_sub_borrow(uint64_t a, uint64_t b, uint64_t borrow, uint64_t &out) {
    const unsigned __int128 s = (unsigned __int128)b + borrow;
    out = a - (uint64_t)s;
    return s > a ? 1 : 0;
}

// hi:lo = a * b.
static inline void precn_mul_wide(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
    const unsigned __int128 p = (unsigned __int128)a * b;
    lo = (uint64_t)p;
    hi = (uint64_t)(p >> 64);
}

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
    // decimal digits, most significant first; other characters are skipped
    // (src/init_basic.cpp:66-92); parsed on the GPU (pg_from_decimal)
    precn_t(std::string o);
    template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
    explicit operator T() const {
        return rsiz == 0 ? (T)0 : (T)a[0];
    }
    explicit operator std::string() const;   // decimal, "0" for zero
    precn_t &operator=(const precn_t &o);
    precn_t &operator=(precn_t &&o);
    ~precn_t();
};

// ---- products (src/mul_basic.cpp, mul_kara.cpp, mul_toom.cpp, mul_fft.cpp,
// mul_ntt.cpp, mul_ssa.cpp) ----------------------------------------------------------
// Each returns the exact product; all of them run on the GPU engine.
precn_t mul_basic(const precn_t &a, const precn_t &b);
precn_t mul_karatsuba(const precn_t &a, const precn_t &b);
precn_t mul_toom(const precn_t &a, const precn_t &b);
precn_t mul_toom23(const precn_t &a, const precn_t &b);
precn_t mul_toom24(const precn_t &a, const precn_t &b);
precn_t mul_toom33(const precn_t &a, const precn_t &b);
precn_t precn_divexact_2(const precn_t &a);   // a / 2 (a even)
precn_t precn_divexact_3(const precn_t &a);   // a / 3 for a multiple of 3 (a 3^-1 mod 2^(64 rsiz))
precn_t mul_fft(const precn_t &a, const precn_t &b);
#if defined(COUNT_FFTS) && COUNT_FFTS
// Rounding statistics of mul_fft (src/mul_fft.cpp:32-37, 359-382), fed from
// the device's largest rounding distance of every FFT product.
extern uint64_t total_fftmuls;
extern uint64_t danger_fftmuls;
extern uint64_t danger_fftmuls_1_4;
extern uint64_t danger_fftmuls_3_8;
extern double max_fft_rounding_error;
#endif
precn_t mul_ntt(const precn_t &a, const precn_t &b);
precn_t mul_ssa(const precn_t &a, const precn_t &b);   // mul_ntt (mul_ssa.cpp:153-158)
// mul_fft plus the largest rounding distance of the call.
precn_t mul_fft_round(const precn_t &a, const precn_t &b, double *max_round);

// ---- addition, subtraction (src/add.cpp, sub.cpp) ---------------------------------------
precn_t operator+(const precn_t &a, const precn_t &b);
precn_t operator-(const precn_t &a, const precn_t &b);   // 0 when a < b (sub.cpp:14)
precn_t add_u64(const precn_t &a, uint64_t b);

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
static inline uint64_t precn_scalar_u64(T b) {
    if (std::is_signed<T>::value && b < 0) std::abort();
    return (uint64_t)b;
}

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator+(const precn_t &a, T b) {
    return add_u64(a, precn_scalar_u64(b));
}

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator+(T a, const precn_t &b) {
    return add_u64(b, precn_scalar_u64(a));
}

// ---- products with the dispatcher (src/mul_basic.cpp:193-284, mul_u32.cpp) -------------
precn_t operator*(const precn_t &a, const precn_t &b);
void mul_into(precn_t &r, const precn_t &a, const precn_t &b);
precn_t mul_u32(const precn_t &a, uint32_t b);
precn_t mul_u64(const precn_t &a, uint64_t b);
precn_t precn_sqr(const precn_t &a);

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator*(const precn_t &a, T b) {
    return mul_u64(a, precn_scalar_u64(b));
}

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator*(T a, const precn_t &b) {
    return mul_u64(b, precn_scalar_u64(a));
}

// ---- division (src/div_basic.cpp, div_dc.cpp, div_newton.cpp) ---------------------------
// Quotients are floors; a zero divisor gives 0 for both quotient and
// remainder (div_basic.cpp:125-133).
precn_t operator/(const precn_t &a, const precn_t &b);
void div_into(precn_t &q, const precn_t &a, const precn_t &b);
void divmod_into(precn_t &q, precn_t &r, const precn_t &a, const precn_t &b);
void divmod_schoolbook_into(precn_t &q, precn_t &r, const precn_t &a, const precn_t &b);
// false (q, r untouched) unless 64 <= b.rsiz, a >= b and a.rsiz <= 2 b.rsiz (div_dc.cpp:92-94)
bool div_dc_into(precn_t &q, precn_t &r, const precn_t &a, const precn_t &b);
precn_t div_schoolbook(const precn_t &a, const precn_t &b);
precn_t mod_schoolbook(const precn_t &a, const precn_t &b);
precn_t div_u32(const precn_t &a, uint32_t b);
precn_t div_u64(const precn_t &a, uint64_t b);
precn_t precn_reciprocal_newton(const precn_t &b, size_t n);
precn_t div_mulinv(const precn_t &a, const precn_t &b);
precn_t mod_mulinv(const precn_t &a, const precn_t &b);
precn_t operator%(const precn_t &a, const precn_t &b);
void mod_into(precn_t &r, const precn_t &a, const precn_t &b);
precn_t mod_u32(const precn_t &a, uint32_t b);
precn_t mod_u64(const precn_t &a, uint64_t b);

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator/(const precn_t &a, T b) {
    return div_u64(a, precn_scalar_u64(b));
}

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator/(T a, const precn_t &b) {
    return precn_t(precn_scalar_u64(a)) / b;
}

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator%(const precn_t &a, T b) {
    return mod_u64(a, precn_scalar_u64(b));
}

template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
precn_t operator%(T a, const precn_t &b) {
    return precn_t(precn_scalar_u64(a)) % b;
}

// ---- shifts, bitwise, increments, comparisons (src/compare_shift.cpp) --------------------
precn_t operator<<(const precn_t &a, size_t b);
precn_t operator>>(const precn_t &a, size_t b);

precn_t operator&(const precn_t &a, const precn_t &b);
precn_t operator|(const precn_t &a, const precn_t &b);
precn_t operator^(const precn_t &a, const precn_t &b);

precn_t operator+(const precn_t &a);   // unary plus: a copy

precn_t operator++(precn_t &a);        // prefix: a + 1
precn_t operator++(precn_t &a, int);   // postfix: returns the old value
precn_t operator--(precn_t &a);        // prefix: a - 1 (0 stays 0, as operator-)
precn_t operator--(precn_t &a, int);

bool operator==(const precn_t &a, const precn_t &b);
bool operator!=(const precn_t &a, const precn_t &b);
bool operator<(const precn_t &a, const precn_t &b);
bool operator>(const precn_t &a, const precn_t &b);
bool operator<=(const precn_t &a, const precn_t &b);
bool operator>=(const precn_t &a, const precn_t &b);

// ---- gcd, square root, base conversion (src/gcd.cpp, sqrt.cpp, base_convert.cpp) ---------
precn_t gcd(const precn_t &a, const precn_t &b);
precn_t precn_sqrt(const precn_t &a);

// Digits of a in base 2..2^32-1, least significant first; out_siz receives
// the count (out may be null to ask for it first; 0 for a == 0).
void precn_base_convert(const precn_t &a, uint32_t base, uint32_t *out, size_t &out_siz);

// ---- signed numbers (prec.hpp:207-289, src/precz.cpp) --------------------------------------
class precz_t {   // sign and precn_t magnitude; zero is never negative
    precn_t mag_;
    bool neg_;

    precz_t(precn_t mag, bool neg);

public:
    precz_t();
    template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
    precz_t(T val) : mag_(), neg_(false) {
        uint64_t bits = (uint64_t)val;
        if (std::is_signed<T>::value && val < 0) {
            neg_ = true;
            bits = (uint64_t)0 - bits;
        }
        mag_ = precn_t(bits);
        if (mag_.rsiz == 0) neg_ = false;
    }
    precz_t(const precn_t &mag);
    precz_t(precn_t &&mag);
    // optional leading whitespace and '-', then the digits precn_t(std::string) reads
    precz_t(std::string val);

    bool is_negative() const;
    bool is_zero() const;
    const precn_t &magnitude() const;

    template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
    explicit operator T() const {
        uint64_t bits = mag_.rsiz == 0 ? 0 : mag_.a[0];
        if (neg_) bits = (uint64_t)0 - bits;
        return (T)bits;
    }
    explicit operator std::string() const;

    precz_t &operator+=(const precz_t &o);
    precz_t &operator-=(const precz_t &o);
    precz_t &operator*=(const precz_t &o);
    precz_t &operator/=(const precz_t &o);
    precz_t &operator%=(const precz_t &o);
    precz_t &operator<<=(size_t bits);
    precz_t &operator>>=(size_t bits);

    friend precz_t operator+(const precz_t &a, const precz_t &b);
    friend precz_t operator-(const precz_t &a, const precz_t &b);
    friend precz_t operator-(const precz_t &a);
    friend precz_t operator+(const precz_t &a);
    friend precz_t operator*(const precz_t &a, const precz_t &b);
    friend precz_t operator/(const precz_t &a, const precz_t &b);
    friend precz_t operator%(const precz_t &a, const precz_t &b);
    friend precz_t operator<<(const precz_t &a, size_t bits);
    friend precz_t operator>>(const precz_t &a, size_t bits);
    friend bool operator==(const precz_t &a, const precz_t &b);
    friend bool operator<(const precz_t &a, const precz_t &b);
    friend precz_t abs(const precz_t &a);
    friend precz_t gcd(const precz_t &a, const precz_t &b);
    friend precz_t precz_sqrt(const precz_t &a);
};

// Quotients truncate toward zero; the remainder takes the dividend's sign
// (precz.cpp:48-54).  Shifts act on the magnitude.
precz_t operator+(const precz_t &a, const precz_t &b);
precz_t operator-(const precz_t &a, const precz_t &b);
precz_t operator-(const precz_t &a);
precz_t operator+(const precz_t &a);
precz_t operator*(const precz_t &a, const precz_t &b);
precz_t operator/(const precz_t &a, const precz_t &b);
precz_t operator%(const precz_t &a, const precz_t &b);
precz_t operator<<(const precz_t &a, size_t bits);
precz_t operator>>(const precz_t &a, size_t bits);

precz_t operator++(precz_t &a);
precz_t operator++(precz_t &a, int);
precz_t operator--(precz_t &a);
precz_t operator--(precz_t &a, int);

bool operator==(const precz_t &a, const precz_t &b);
bool operator!=(const precz_t &a, const precz_t &b);
bool operator<(const precz_t &a, const precz_t &b);
bool operator>(const precz_t &a, const precz_t &b);
bool operator<=(const precz_t &a, const precz_t &b);
bool operator>=(const precz_t &a, const precz_t &b);

precz_t abs(const precz_t &a);
precz_t gcd(const precz_t &a, const precz_t &b);
precz_t precz_sqrt(const precz_t &a);   // zero for negative inputs
