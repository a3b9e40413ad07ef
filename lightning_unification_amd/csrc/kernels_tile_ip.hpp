// In-place complex128 updates of register-resident tile amplitudes.
//
// Each update is one block of inline assembly whose results overwrite the
// registers of the amplitudes they replace (tied "+v" operands), using two or
// four early-clobber temporaries.  Written in C++, the compiler computes an
// op's new amplitudes into a second register set (all products first, for
// latency) and copies them back at the end of the op, because a tied
// v_fmac_f64 result cannot share the register of an input that is still read
// afterwards.  In the tile-pass interpreter that kept a second copy of the
// tile live and added a copy per amplitude per op.  Here every op keeps the
// tile in one fixed register set (with the interpreter's other measures,
// 154 VGPRs for the complex128 2^10-amplitude tile, from 214).
//
// Operand rules of the VOP3 encodings used (gfx9 family): at most one SGPR
// source per instruction (ip_pair_s takes a uniform matrix from SGPRs and
// reads exactly one matrix element per instruction; all else is VGPRs); a
// source negation is the "-" modifier.  Only FP64 multiply and FMA appear;
// dependent instructions are interlocked by the hardware.
#pragma once

namespace plamd {

// x <- a x
__device__ __forceinline__ void ip_scale(double &xr, double &xi, double ar, double ai) {
    double t, u;
    asm("v_mul_f64 %[t], %[ai], %[xi]\n\t"
        "v_mul_f64 %[u], %[ai], %[xr]\n\t"
        "v_fma_f64 %[xr], %[ar], %[xr], -%[t]\n\t"
        "v_fma_f64 %[xi], %[ar], %[xi], %[u]"
        : [xr] "+v"(xr), [xi] "+v"(xi), [t] "=&v"(t), [u] "=&v"(u)
        : [ar] "v"(ar), [ai] "v"(ai));
}

// (x, y) <- (m00 x + m01 y, m10 x + m11 y), the matrix uniform (SGPRs)
__device__ __forceinline__ void ip_pair_s(double &xr, double &xi, double &yr, double &yi, double m00r,
                                          double m00i, double m01r, double m01i, double m10r, double m10i,
                                          double m11r, double m11i) {
    double tr, ti, ur, ui;
    asm(// u = m10 x
        "v_mul_f64 %[ur], %[m10r], %[xr]\n\t"
        "v_fma_f64 %[ur], -%[m10i], %[xi], %[ur]\n\t"
        "v_mul_f64 %[ui], %[m10r], %[xi]\n\t"
        "v_fma_f64 %[ui], %[m10i], %[xr], %[ui]\n\t"
        // t = m01 y
        "v_mul_f64 %[tr], %[m01r], %[yr]\n\t"
        "v_fma_f64 %[tr], -%[m01i], %[yi], %[tr]\n\t"
        "v_mul_f64 %[ti], %[m01r], %[yi]\n\t"
        "v_fma_f64 %[ti], %[m01i], %[yr], %[ti]\n\t"
        // x = m00 x + t
        "v_fma_f64 %[tr], -%[m00i], %[xi], %[tr]\n\t"
        "v_fma_f64 %[ti], %[m00i], %[xr], %[ti]\n\t"
        "v_fma_f64 %[xr], %[m00r], %[xr], %[tr]\n\t"
        "v_fma_f64 %[xi], %[m00r], %[xi], %[ti]\n\t"
        // y = m11 y + u
        "v_fma_f64 %[ur], -%[m11i], %[yi], %[ur]\n\t"
        "v_fma_f64 %[ui], %[m11i], %[yr], %[ui]\n\t"
        "v_fma_f64 %[yr], %[m11r], %[yr], %[ur]\n\t"
        "v_fma_f64 %[yi], %[m11r], %[yi], %[ui]"
        : [xr] "+v"(xr), [xi] "+v"(xi), [yr] "+v"(yr), [yi] "+v"(yi), [tr] "=&v"(tr), [ti] "=&v"(ti),
          [ur] "=&v"(ur), [ui] "=&v"(ui)
        : [m00r] "s"(m00r), [m00i] "s"(m00i), [m01r] "s"(m01r), [m01i] "s"(m01i), [m10r] "s"(m10r),
          [m10i] "s"(m10i), [m11r] "s"(m11r), [m11i] "s"(m11i));
}

// (x, y) <- (m00 x + m01 y, m10 x + m11 y), the matrix per lane (VGPRs)
__device__ __forceinline__ void ip_pair_v(double &xr, double &xi, double &yr, double &yi, double m00r,
                                          double m00i, double m01r, double m01i, double m10r, double m10i,
                                          double m11r, double m11i) {
    double tr, ti, ur, ui;
    asm(// u = m10 x
        "v_mul_f64 %[ur], %[m10r], %[xr]\n\t"
        "v_fma_f64 %[ur], -%[m10i], %[xi], %[ur]\n\t"
        "v_mul_f64 %[ui], %[m10r], %[xi]\n\t"
        "v_fma_f64 %[ui], %[m10i], %[xr], %[ui]\n\t"
        // t = m01 y
        "v_mul_f64 %[tr], %[m01r], %[yr]\n\t"
        "v_fma_f64 %[tr], -%[m01i], %[yi], %[tr]\n\t"
        "v_mul_f64 %[ti], %[m01r], %[yi]\n\t"
        "v_fma_f64 %[ti], %[m01i], %[yr], %[ti]\n\t"
        // x = m00 x + t
        "v_fma_f64 %[tr], -%[m00i], %[xi], %[tr]\n\t"
        "v_fma_f64 %[ti], %[m00i], %[xr], %[ti]\n\t"
        "v_fma_f64 %[xr], %[m00r], %[xr], %[tr]\n\t"
        "v_fma_f64 %[xi], %[m00r], %[xi], %[ti]\n\t"
        // y = m11 y + u
        "v_fma_f64 %[ur], -%[m11i], %[yi], %[ur]\n\t"
        "v_fma_f64 %[ui], %[m11i], %[yr], %[ui]\n\t"
        "v_fma_f64 %[yr], %[m11r], %[yr], %[ur]\n\t"
        "v_fma_f64 %[yi], %[m11r], %[yi], %[ui]"
        : [xr] "+v"(xr), [xi] "+v"(xi), [yr] "+v"(yr), [yi] "+v"(yi), [tr] "=&v"(tr), [ti] "=&v"(ti),
          [ur] "=&v"(ur), [ui] "=&v"(ui)
        : [m00r] "v"(m00r), [m00i] "v"(m00i), [m01r] "v"(m01r), [m01i] "v"(m01i), [m10r] "v"(m10r),
          [m10i] "v"(m10i), [m11r] "v"(m11r), [m11i] "v"(m11i));
}

// x <- a x + b o (o: the partner amplitude, read only; a, b per lane)
__device__ __forceinline__ void ip_lane(double &xr, double &xi, double orr, double oi, double ar, double ai,
                                        double br, double bi) {
    double tr, ti;
    asm("v_mul_f64 %[tr], %[br], %[pr]\n\t"
        "v_fma_f64 %[tr], -%[bi], %[oi], %[tr]\n\t"
        "v_mul_f64 %[ti], %[br], %[oi]\n\t"
        "v_fma_f64 %[ti], %[bi], %[pr], %[ti]\n\t"
        "v_fma_f64 %[tr], -%[ai], %[xi], %[tr]\n\t"
        "v_fma_f64 %[ti], %[ai], %[xr], %[ti]\n\t"
        "v_fma_f64 %[xr], %[ar], %[xr], %[tr]\n\t"
        "v_fma_f64 %[xi], %[ar], %[xi], %[ti]"
        : [xr] "+v"(xr), [xi] "+v"(xi), [tr] "=&v"(tr), [ti] "=&v"(ti)
        : [pr] "v"(orr), [oi] "v"(oi), [ar] "v"(ar), [ai] "v"(ai), [br] "v"(br), [bi] "v"(bi));
}

// (a0, a1) <- (w0 a0 + w1 a1 + w2 p0 + w3 p1, w4 a0 + w5 a1 + w6 p0 + w7 p1)
// with p0, p1 the partner lane's amplitudes; w per lane, w[2c], w[2c+1] =
// real and imaginary part of w_c
__device__ __forceinline__ void ip_mix1(double &a0r, double &a0i, double &a1r, double &a1i, double p0r,
                                        double p0i, double p1r, double p1i, const double (&w)[16]) {
    double tr, ti, ur, ui;
    asm(// t = w2 p0 + w3 p1 + w1 a1
        "v_mul_f64 %[tr], %[w2r], %[p0r]\n\t"
        "v_fma_f64 %[tr], -%[w2i], %[p0i], %[tr]\n\t"
        "v_mul_f64 %[ti], %[w2r], %[p0i]\n\t"
        "v_fma_f64 %[ti], %[w2i], %[p0r], %[ti]\n\t"
        "v_fma_f64 %[tr], %[w3r], %[p1r], %[tr]\n\t"
        "v_fma_f64 %[tr], -%[w3i], %[p1i], %[tr]\n\t"
        "v_fma_f64 %[ti], %[w3r], %[p1i], %[ti]\n\t"
        "v_fma_f64 %[ti], %[w3i], %[p1r], %[ti]\n\t"
        "v_fma_f64 %[tr], %[w1r], %[a1r], %[tr]\n\t"
        "v_fma_f64 %[tr], -%[w1i], %[a1i], %[tr]\n\t"
        "v_fma_f64 %[ti], %[w1r], %[a1i], %[ti]\n\t"
        "v_fma_f64 %[ti], %[w1i], %[a1r], %[ti]\n\t"
        // u = w6 p0 + w7 p1 + w4 a0
        "v_mul_f64 %[ur], %[w6r], %[p0r]\n\t"
        "v_fma_f64 %[ur], -%[w6i], %[p0i], %[ur]\n\t"
        "v_mul_f64 %[ui], %[w6r], %[p0i]\n\t"
        "v_fma_f64 %[ui], %[w6i], %[p0r], %[ui]\n\t"
        "v_fma_f64 %[ur], %[w7r], %[p1r], %[ur]\n\t"
        "v_fma_f64 %[ur], -%[w7i], %[p1i], %[ur]\n\t"
        "v_fma_f64 %[ui], %[w7r], %[p1i], %[ui]\n\t"
        "v_fma_f64 %[ui], %[w7i], %[p1r], %[ui]\n\t"
        "v_fma_f64 %[ur], %[w4r], %[a0r], %[ur]\n\t"
        "v_fma_f64 %[ur], -%[w4i], %[a0i], %[ur]\n\t"
        "v_fma_f64 %[ui], %[w4r], %[a0i], %[ui]\n\t"
        "v_fma_f64 %[ui], %[w4i], %[a0r], %[ui]\n\t"
        // a0 = w0 a0 + t
        "v_fma_f64 %[tr], -%[w0i], %[a0i], %[tr]\n\t"
        "v_fma_f64 %[ti], %[w0i], %[a0r], %[ti]\n\t"
        "v_fma_f64 %[a0r], %[w0r], %[a0r], %[tr]\n\t"
        "v_fma_f64 %[a0i], %[w0r], %[a0i], %[ti]\n\t"
        // a1 = w5 a1 + u
        "v_fma_f64 %[ur], -%[w5i], %[a1i], %[ur]\n\t"
        "v_fma_f64 %[ui], %[w5i], %[a1r], %[ui]\n\t"
        "v_fma_f64 %[a1r], %[w5r], %[a1r], %[ur]\n\t"
        "v_fma_f64 %[a1i], %[w5r], %[a1i], %[ui]"
        : [a0r] "+v"(a0r), [a0i] "+v"(a0i), [a1r] "+v"(a1r), [a1i] "+v"(a1i), [tr] "=&v"(tr), [ti] "=&v"(ti),
          [ur] "=&v"(ur), [ui] "=&v"(ui)
        : [p0r] "v"(p0r), [p0i] "v"(p0i), [p1r] "v"(p1r), [p1i] "v"(p1i), [w0r] "v"(w[0]), [w0i] "v"(w[1]),
          [w1r] "v"(w[2]), [w1i] "v"(w[3]), [w2r] "v"(w[4]), [w2i] "v"(w[5]), [w3r] "v"(w[6]),
          [w3i] "v"(w[7]), [w4r] "v"(w[8]), [w4i] "v"(w[9]), [w5r] "v"(w[10]), [w5i] "v"(w[11]),
          [w6r] "v"(w[12]), [w6i] "v"(w[13]), [w7r] "v"(w[14]), [w7i] "v"(w[15]));
}

// x <- w0 x + w1 s1 + w2 s2 + w3 s3 (s: partner amplitudes; w per lane)
__device__ __forceinline__ void ip_mix2(double &xr, double &xi, double s1r, double s1i, double s2r, double s2i,
                                        double s3r, double s3i, const double (&w)[8]) {
    double tr, ti;
    asm("v_mul_f64 %[tr], %[w1r], %[s1r]\n\t"
        "v_fma_f64 %[tr], -%[w1i], %[s1i], %[tr]\n\t"
        "v_mul_f64 %[ti], %[w1r], %[s1i]\n\t"
        "v_fma_f64 %[ti], %[w1i], %[s1r], %[ti]\n\t"
        "v_fma_f64 %[tr], %[w2r], %[s2r], %[tr]\n\t"
        "v_fma_f64 %[tr], -%[w2i], %[s2i], %[tr]\n\t"
        "v_fma_f64 %[ti], %[w2r], %[s2i], %[ti]\n\t"
        "v_fma_f64 %[ti], %[w2i], %[s2r], %[ti]\n\t"
        "v_fma_f64 %[tr], %[w3r], %[s3r], %[tr]\n\t"
        "v_fma_f64 %[tr], -%[w3i], %[s3i], %[tr]\n\t"
        "v_fma_f64 %[ti], %[w3r], %[s3i], %[ti]\n\t"
        "v_fma_f64 %[ti], %[w3i], %[s3r], %[ti]\n\t"
        "v_fma_f64 %[tr], -%[w0i], %[xi], %[tr]\n\t"
        "v_fma_f64 %[ti], %[w0i], %[xr], %[ti]\n\t"
        "v_fma_f64 %[xr], %[w0r], %[xr], %[tr]\n\t"
        "v_fma_f64 %[xi], %[w0r], %[xi], %[ti]"
        : [xr] "+v"(xr), [xi] "+v"(xi), [tr] "=&v"(tr), [ti] "=&v"(ti)
        : [s1r] "v"(s1r), [s1i] "v"(s1i), [s2r] "v"(s2r), [s2i] "v"(s2i), [s3r] "v"(s3r), [s3i] "v"(s3i),
          [w0r] "v"(w[0]), [w0i] "v"(w[1]), [w1r] "v"(w[2]), [w1i] "v"(w[3]), [w2r] "v"(w[4]),
          [w2i] "v"(w[5]), [w3r] "v"(w[6]), [w3i] "v"(w[7]));
}

} // namespace plamd
