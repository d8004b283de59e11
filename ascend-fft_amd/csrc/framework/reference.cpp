// FP64 reference transforms of the C++ API (include/butterfly/reference.hpp),
// used by Plan::measure's acceptance check.  Implemented as an FP64 radix-2
// Stockham autosort (ping-pong buffers, no bit reversal) with twiddles from
// exact integer angle reduction -- a different algorithm from the device
// kernels' radix-16 passes, so the check is not the kernel re-run on the CPU.
#include <cmath>
#include <complex>
#include <vector>

#include "butterfly/reference.hpp"

namespace bfly {
namespace {

using cd = std::complex<double>;

// y = DFT_n(x) with e^{sign 2 pi i jk / n}; n a power of two.
void stockham64(std::vector<cd>& x, int sign) {
    const size_t n = x.size();
    if (n < 2) return;
    std::vector<cd> y(n), w(n / 2);
    for (size_t m = 0; m < n / 2; m++) {
        const double a = sign * 2.0 * M_PI * (double)m / (double)n;
        w[m] = cd(std::cos(a), std::sin(a));
    }
    // Pass with l = current half-length of combined blocks:
    //   y[j*2l + k] = x[j*l + k] + w * x[j*l + k + n/2]
    //   y[j*2l + k + l] = x[j*l + k] - w * x[j*l + k + n/2],  w = W_{2l}^k
    for (size_t l = 1; l < n; l <<= 1) {
        const size_t blocks = n / (2 * l);
        for (size_t j = 0; j < blocks; j++) {
            for (size_t k = 0; k < l; k++) {
                const cd a = x[j * l + k];
                const cd b = x[j * l + k + n / 2] * w[k * blocks];
                y[j * 2 * l + k] = a + b;
                y[j * 2 * l + k + l] = a - b;
            }
        }
        x.swap(y);
    }
}

}  // namespace

float patternAt(size_t i) { return (float)((i * 2654435761u) % 1000) / 1000.f - 0.5f; }

void refFftF32(const float* in, float* out, uint32_t n) {
    if (!n) return;
    std::vector<cd> x(n);
    for (uint32_t j = 0; j < n; j++) x[j] = cd(in[2 * j], in[2 * j + 1]);
    stockham64(x, -1);
    for (uint32_t k = 0; k < n; k++) {
        out[2 * k] = (float)x[k].real();
        out[2 * k + 1] = (float)x[k].imag();
    }
}

void refR2CF32(const float* in, float* out, uint32_t n) {
    if (n < 2) return;
    std::vector<cd> x(n);
    for (uint32_t j = 0; j < n; j++) x[j] = cd(in[j], 0.0);
    stockham64(x, -1);
    for (uint32_t k = 0; k <= n / 2; k++) {
        out[2 * k] = (float)x[k].real();
        out[2 * k + 1] = (float)x[k].imag();
    }
}

void refC2RF32(const float* in, float* out, uint32_t n) {
    if (n < 2) return;
    const uint32_t m = n / 2;
    std::vector<cd> X(n);
    X[0] = cd(in[0], 0.0);
    X[m] = cd(in[2 * m], 0.0);
    for (uint32_t k = 1; k < m; k++) {
        X[k] = cd(in[2 * k], in[2 * k + 1]);
        X[n - k] = std::conj(X[k]);
    }
    stockham64(X, +1);
    for (uint32_t j = 0; j < n; j++) out[j] = (float)(X[j].real() / (double)n);
}

}  // namespace bfly
