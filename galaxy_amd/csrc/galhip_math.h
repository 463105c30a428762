// galhip_math.h -- the logarithm and arc tangent of the cell-index
// computation, written with IEEE +, -, *, / and exact scalings only.
//
// The reference evaluates u = log(1 + r) / alpha in FP32 and
// t = atan2(dble(y), dble(x)) * na / (2 pi) in FP64 rounded to FP32
// (SRC/lib15/uofgr.f:24, weight.f:76-77).  Reproducing its cell indices
// and weights bit for bit needs the correctly rounded FP32 values, i.e. the
// logarithm and arc tangent to about double precision.  These routines
// reach that (error below 1.5 ulp of double, measured against 64-bit
// extended references) with a short table reduction instead of the general
// device library routines, and because they use only correctly rounded
// basic operations (compiled without contraction), a host build returns the
// same bits as the device build, so their accuracy is tested on the CPU.
//
// Tables (galaxy_amd/csrc/gen_math_tables.py, 60-digit decimal):
//   ln(1 + j/32) and atan(j/16) as hi + lo pairs; ln 2, pi/2, pi likewise.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define GH_HD __host__ __device__ __forceinline__
#else
#define GH_HD static inline
#endif

// The tables, in one array so that a kernel can stage it in LDS:
// [0, 32) ln(1 + j/32) hi, [32, 64) its lo, [64, 81) atan(j/16) hi,
// [81, 98) its lo.
#define GH_MATH_TAB_N 98
#define GH_LOGC_HI 0
#define GH_LOGC_LO 32
#define GH_ATANC_HI 64
#define GH_ATANC_LO 81
static constexpr double GH_MATH_TAB[GH_MATH_TAB_N] = {
    // LOGC_HI
    0x0.0p+0, 0x1.f829b0e800000p-6,
    0x1.f0a30c0200000p-5, 0x1.6f0d28ae00000p-4,
    0x1.e27076e200000p-4, 0x1.29552f8200000p-3,
    0x1.5ff3070a00000p-3, 0x1.9525a9d000000p-3,
    0x1.c8ff7c7a00000p-3, 0x1.fb9186d600000p-3,
    0x1.1675caba00000p-2, 0x1.2e8e2bae00000p-2,
    0x1.4618bc2200000p-2, 0x1.5d1bdbf600000p-2,
    0x1.739d7f6c00000p-2, 0x1.89a3386c00000p-2,
    0x1.9f323ecc00000p-2, 0x1.b44f77bc00000p-2,
    0x1.c8ff7c7a00000p-2, 0x1.dd46a04c00000p-2,
    0x1.f128f5fa00000p-2, 0x1.02552a5a00000p-1,
    0x1.0be72e4200000p-1, 0x1.154c3d3000000p-1,
    0x1.1e85f5e800000p-1, 0x1.2795e12800000p-1,
    0x1.307d733400000p-1, 0x1.393e0d3600000p-1,
    0x1.41d8fe8400000p-1, 0x1.4a4f85dc00000p-1,
    0x1.52a2d26600000p-1, 0x1.5ad404c400000p-1,
    // LOGC_LO
    0x0.0p+0, -0x1.f33fecc1c0fb1p-40,
    -0x1.d3ab33d066d1dp-38, 0x1.5ad2e6f9266e8p-38,
    0x1.5e5cbd3d51000p-37, -0x1.5b967f4471dfcp-44,
    0x1.e4f4f21cf8828p-37, -0x1.75297137d9f16p-36,
    -0x1.597794f689f84p-37, -0x1.c1d572aab993dp-39,
    0x1.74c1c07398fabp-35, 0x1.1d309c2cc91a8p-38,
    -0x1.d09ec17a42642p-37, -0x1.fd8d6bdc9c7c2p-36,
    -0x1.0bfe58c76ceb0p-36, 0x1.425ab5a718811p-38,
    -0x1.9ed03525ca264p-40, 0x1.91ec5197ddb56p-35,
    -0x1.597794f689f84p-36, 0x1.c4a0bee626a4ap-38,
    0x1.e0dd966b90762p-35, 0x1.743fb1a71a576p-35,
    0x1.4aa0ada625eedp-35, -0x1.6542cace198b9p-34,
    -0x1.f7e5f84274cb4p-34, 0x1.36235d6f07e7bp-34,
    0x1.e217c3f6b2144p-34, -0x1.3abccac777b40p-34,
    0x1.9cab99192f30cp-35, -0x1.f8289fbb08171p-34,
    -0x1.0e9544620dd44p-35, -0x1.4c1a609acaab4p-34,
    // ATANC_HI
    0x0.0p+0, 0x1.ff55bb72cfdeap-5,
    0x1.fd5ba9aac2f6ep-4, 0x1.7b97b4bce5b02p-3,
    0x1.f5b75f92c80ddp-3, 0x1.362773707ebccp-2,
    0x1.6f61941e4def1p-2, 0x1.a64eec3cc23fdp-2,
    0x1.dac670561bb4fp-2, 0x1.0657e94db30d0p-1,
    0x1.1e00babdefeb4p-1, 0x1.345f01cce37bbp-1,
    0x1.4978fa3269ee1p-1, 0x1.5d58987169b18p-1,
    0x1.700a7c5784634p-1, 0x1.819d0b7158a4dp-1,
    0x1.921fb54442d18p-1,
    // ATANC_LO
    0x0.0p+0, -0x1.c934d86d23f1dp-60,
    -0x1.cd37686760c17p-59, 0x1.347b0b4f881cap-58,
    0x1.8ab6e3cf7afbdp-57, -0x1.963a544b672d8p-57,
    -0x1.c63aae6f6e918p-56, -0x1.24dec1b50b7ffp-56,
    0x1.a2b7f222f65e2p-56, -0x1.d5b495f6349e6p-56,
    -0x1.928df287a668fp-58, 0x1.1021137c71102p-55,
    0x1.2419a87f2a458p-56, 0x1.0028e4bc5e7cap-57,
    -0x1.8c34d25aadef6p-56, -0x1.bf76229d3b917p-56,
    0x1.1a62633145c07p-55,
};

// ln(v) for v >= 1.  v = m 2^e, m in [1, 2); c = 1 + j/32 the table point
// at or below m; ln m = ln c + 2 atanh(s), s = (m - c)/(m + c), |s| < 1/64.
GH_HD double gh_log_ge1(double v, const double *tab = GH_MATH_TAB) {
    const double LN2_HI = 0x1.62e42ff000000p-1, LN2_LO = -0x1.718432a1b0e26p-35;
    int e;
    double m = frexp(v, &e);          // m in [0.5, 1)
    m = m * 2.0;
    e = e - 1;
    int j = (int)((m - 1.0) * 32.0);
    double c = 1.0 + (double)j * 0.03125;
    double s = (m - c) / (m + c);
    double z = s * s;
    double p = z * (1.0 / 3.0 + z * (1.0 / 5.0 + z * (1.0 / 7.0 + z * (1.0 / 9.0))));
    double two_s = s + s;
    double fe = (double)e;
    double hi = fe * LN2_HI + tab[GH_LOGC_HI + j];
    double lo = (fe * LN2_LO + tab[GH_LOGC_LO + j]) + two_s * p;
    return hi + (two_s + lo);
}

// atan2(y, x) with IEEE signed-zero conventions for finite arguments.
// a = min/max of |y|, |x| in [0, 1]; c = j/16 the nearest table point;
// atan a = atan c + atan z, z = (a - c)/(1 + a c), |z| <= 1/32.
GH_HD double gh_atan2(double y, double x, const double *tab = GH_MATH_TAB) {
    const double PIO2_HI = 0x1.921fb54442d18p+0, PIO2_LO = 0x1.1a62633145c07p-54;
    const double PI_HI = 0x1.921fb54442d18p+1, PI_LO = 0x1.1a62633145c07p-53;
    double ax = fabs(x), ay = fabs(y);
    bool neg_y = signbit(y), neg_x = signbit(x);
    double r;
    if (ax == 0.0 && ay == 0.0) {
        r = neg_x ? PI_HI : 0.0;
    } else {
        bool swap = ay > ax;
        // one quotient of the selected operands (a select of two quotients
        // would evaluate both divisions)
        const double num = swap ? ax : ay, den = swap ? ay : ax;
        double a = num / den;
        int j = (int)(a * 16.0 + 0.5);
        double c = (double)j * 0.0625;
        double z = (a - c) / (1.0 + a * c);
        double z2 = z * z;
        double q = z2 * (-1.0 / 3.0 + z2 * (1.0 / 5.0 + z2 * (-1.0 / 7.0 +
                   z2 * (1.0 / 9.0 + z2 * (-1.0 / 11.0)))));
        // atan a = hi + (lo + z + z q)
        double t = tab[GH_ATANC_HI + j] + (tab[GH_ATANC_LO + j] + (z + z * q));
        if (swap) t = PIO2_HI - (t - PIO2_LO);
        if (neg_x) t = PI_HI - (t - PI_LO);
        r = t;
    }
    return neg_y ? -r : r;
}
