// GF(2) polynomial machinery for MT19937 jump-ahead (host side).
//
// MT19937 is linear over GF(2): its 19937-bit live state s_n = (msb(y_n), y_{n+1},
// ..., y_{n+623}) evolves by a matrix A whose characteristic polynomial phi has
// degree 19937. For any J, A^J = c_J(A) with c_J = x^J mod phi (Cayley-Hamilton),
// so the state J words ahead is XOR_{i: c_J[i]=1} s_i, i.e. word w of the jumped
// window equals XOR over the set coefficients i of y_{i+w} (the jump method of
// Haramoto, Matsumoto, Nishimura, L'Ecuyer, Panneton 2008, applied in its
// "combine shifted windows" form, which a GPU evaluates word-parallel).
//
// phi is obtained once per process by Berlekamp-Massey on an output bit stream;
// the jump polynomials x^(2^e) mod phi by repeated squaring. Both are independent
// of the seed, so they are computed once and cached.

#include "fm_internal.hpp"

#include <mutex>
#include <cstring>

namespace fm {

namespace {

constexpr int kDeg = 19937;                 // degree of phi
constexpr int kWords = (kDeg + 63) / 64;    // 312 words hold a reduced polynomial

struct Gf2Tables {
    std::vector<uint64_t> phi;              // kDeg+1 bits, little-endian bit order
    std::vector<std::vector<uint64_t>> pow2; // pow2[e] = x^(2^e) mod phi
};

std::once_flag g_once;
Gf2Tables* g_tables = nullptr;

// Reference MT19937 (mt19937ar) for the bit stream fed to Berlekamp-Massey.
struct SeqMT {
    uint32_t mt[624]; int mti;
    explicit SeqMT(uint32_t seed) {
        mt[0] = seed;
        for (int i = 1; i < 624; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
        mti = 624;
    }
    uint32_t next() {
        if (mti >= 624) {
            for (int k = 0; k < 624; k++) {
                uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            mti = 0;
        }
        uint32_t y = mt[mti++];
        y ^= y >> 11; y ^= (y << 7) & 0x9d2c5680u; y ^= (y << 15) & 0xefc60000u; y ^= y >> 18;
        return y;
    }
};

inline int getbit(const std::vector<uint64_t>& v, int i) { return (int)((v[(size_t)i >> 6] >> (i & 63)) & 1u); }
inline void flipbit(std::vector<uint64_t>& v, int i) { v[(size_t)i >> 6] ^= (uint64_t)1 << (i & 63); }

// Berlekamp-Massey over GF(2). Returns the connection polynomial C (bits c_0..c_L).
std::vector<uint64_t> berlekamp_massey(const std::vector<uint8_t>& s, int& Lout) {
    const int N = (int)s.size();
    const int W = (N + 64) / 64 + 1;
    std::vector<uint64_t> C(W, 0), B(W, 0), T;
    C[0] = B[0] = 1;
    int L = 0, m = 1;
    // reversed sequence so that sum_{i=1..L} c_i s[n-i] becomes a word-wise AND
    std::vector<uint64_t> rs((N + 63) / 64 + 2, 0);
    for (int i = 0; i < N; i++) if (s[i]) rs[(size_t)(N - 1 - i) >> 6] |= (uint64_t)1 << ((N - 1 - i) & 63);
    auto rbits = [&](int start) -> uint64_t {   // 64 bits of rs starting at bit 'start'
        int w = start >> 6, b = start & 63;
        uint64_t lo = rs[(size_t)w] >> b;
        uint64_t hi = (b && (size_t)(w + 1) < rs.size()) ? (rs[(size_t)w + 1] << (64 - b)) : 0;
        return lo | hi;
    };
    for (int n = 0; n < N; n++) {
        // d = s[n] ^ sum_{i=1..L} c_i s[n-i];  s[n-i] = rs[N-1-n+i]
        uint64_t acc = 0;
        const int base = N - 1 - n;              // rs index of s[n]
        for (int i0 = 0; i0 <= L; i0 += 64) {
            uint64_t cw = C[(size_t)i0 >> 6];
            uint64_t sw = rbits(base + i0);
            if (L - i0 < 63) { int keep = L - i0 + 1; cw &= (keep >= 64) ? ~0ull : ((1ull << keep) - 1); }
            acc ^= cw & sw;
        }
        int d = __builtin_parityll(acc);        // includes c_0 s[n] = s[n]
        if (!d) { m++; continue; }
        if (2 * L <= n) {
            T = C;
            // C ^= B << m
            int ws = m >> 6, bs = m & 63;
            for (int w = W - 1; w >= 0; w--) {
                uint64_t v = 0;
                int src = w - ws;
                if (src >= 0) { v = B[(size_t)src] << bs; if (bs && src - 1 >= 0) v |= B[(size_t)src - 1] >> (64 - bs); }
                C[(size_t)w] ^= v;
            }
            L = n + 1 - L; B = T; m = 1;
        } else {
            int ws = m >> 6, bs = m & 63;
            for (int w = W - 1; w >= 0; w--) {
                uint64_t v = 0;
                int src = w - ws;
                if (src >= 0) { v = B[(size_t)src] << bs; if (bs && src - 1 >= 0) v |= B[(size_t)src - 1] >> (64 - bs); }
                C[(size_t)w] ^= v;
            }
            m++;
        }
    }
    Lout = L;
    return C;
}

// Reduce a polynomial of degree < 2*kDeg modulo phi, in place; result in the low kWords.
void reduce_mod_phi(std::vector<uint64_t>& p, const std::vector<uint64_t>& phi) {
    const int top = (int)p.size() * 64 - 1;
    for (int d = top; d >= kDeg; d--) {
        if (!getbit(p, d)) continue;
        const int sh = d - kDeg;                 // XOR phi << sh (clears bit d)
        const int ws = sh >> 6, bs = sh & 63;
        const int nw = (int)phi.size();
        for (int w = 0; w < nw; w++) {
            uint64_t v = phi[(size_t)w];
            if (!v) continue;
            size_t dst = (size_t)(w + ws);
            if (dst < p.size()) p[dst] ^= v << bs;
            if (bs && dst + 1 < p.size()) p[dst + 1] ^= v >> (64 - bs);
        }
    }
    p.resize(kWords);
}

std::vector<uint64_t> square_mod_phi(const std::vector<uint64_t>& a, const std::vector<uint64_t>& phi) {
    std::vector<uint64_t> sq((size_t)kWords * 2 + 1, 0);
    for (int i = 0; i < kDeg; i++) if (getbit(a, i)) flipbit(sq, 2 * i);
    reduce_mod_phi(sq, phi);
    return sq;
}

void build_tables() {
    auto* t = new Gf2Tables();
    // bit stream: least significant bit of successive outputs (a linear functional of the state)
    const int N = 2 * kDeg + 128;
    std::vector<uint8_t> s((size_t)N);
    SeqMT g(5489u);
    for (int i = 0; i < N; i++) s[(size_t)i] = (uint8_t)(g.next() & 1u);
    int L = 0;
    std::vector<uint64_t> C = berlekamp_massey(s, L);
    if (L != kDeg) throw FmError(FM_ERR_INTERNAL, "MT19937 characteristic polynomial has unexpected degree");
    // phi(x) = x^L C(1/x): coefficient of x^(L-i) is c_i
    t->phi.assign((size_t)(kDeg / 64 + 1), 0);
    for (int i = 0; i <= L; i++) if (getbit(C, i)) flipbit(t->phi, L - i);
    // x^(2^e) mod phi
    std::vector<uint64_t> x((size_t)kWords, 0);
    x[0] = 2;                                     // the polynomial x
    t->pow2.push_back(x);
    for (int e = 1; e <= 40; e++) t->pow2.push_back(square_mod_phi(t->pow2.back(), t->phi));
    g_tables = t;
}

const Gf2Tables& tables() {
    std::call_once(g_once, build_tables);
    return *g_tables;
}

} // namespace

int mt_jump_degree() { return kDeg; }
int mt_jump_words() { return kWords; }

const std::vector<uint64_t>& mt_jump_poly_pow2(int e) {
    const auto& t = tables();
    if (e < 0 || e >= (int)t.pow2.size()) throw FmError(FM_ERR_INVALID_ARGUMENT, "jump exponent out of range");
    return t.pow2[(size_t)e];
}

const std::vector<uint64_t>& mt_char_poly() { return tables().phi; }

// Host-side jump of one window (used by tests of the library and for tiny streams):
// window_J[w] = XOR_{i: c[i]} y[i+w] with y generated from the source window.
void mt_jump_window_host(const uint32_t* src624, const std::vector<uint64_t>& c, uint32_t* dst624) {
    std::vector<uint32_t> y((size_t)kDeg + 624);
    std::memcpy(y.data(), src624, 624 * sizeof(uint32_t));
    for (int n = 0; n + 624 < (int)y.size(); n++) {
        uint32_t v = (y[(size_t)n] & 0x80000000u) | (y[(size_t)n + 1] & 0x7fffffffu);
        y[(size_t)n + 624] = y[(size_t)n + 397] ^ (v >> 1) ^ ((v & 1u) ? 0x9908b0dfu : 0u);
    }
    for (int w = 0; w < 624; w++) dst624[w] = 0;
    for (int i = 0; i < kDeg; i++) {
        if (!((c[(size_t)i >> 6] >> (i & 63)) & 1u)) continue;
        for (int w = 0; w < 624; w++) dst624[w] ^= y[(size_t)i + (size_t)w];
    }
}

// Advance a window by `words` outputs: the product of the jumps x^(2^e) of the set bits
// (jumps commute), each applied on the host.
void mt_jump_window_by(uint32_t* window624, uint64_t words) {
    uint32_t tmp[624];
    for (int e = 0; words != 0; e++, words >>= 1) {
        if (!(words & 1u)) continue;
        mt_jump_window_host(window624, mt_jump_poly_pow2(e), tmp);
        std::memcpy(window624, tmp, sizeof(tmp));
    }
}

} // namespace fm
