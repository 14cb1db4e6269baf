// Host check of the Stockham passes of parties-vof-ibm_amd/csrc/pv_fftk.h:
// the mixed-radix transform (sg_load / sg_store in the pass order of sg_fft)
// and the power-of-two one (sk_load / sk_store) against a direct DFT.
// Reads M from argv, prints the largest relative error of forward and
// inverse transforms of a fixed pseudo-random sequence.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "../../parties-vof-ibm_amd/csrc/pv_fftk.h"

template <int R, bool INV>
static void host_pass_g(std::vector<double> &sr, std::vector<double> &si, int M, int Ns, const double *tw)
{
    const int nbf = M / R;
    std::vector<double> vr((size_t)nbf * R), vi((size_t)nbf * R);
    for (int bb = 0; bb < nbf; bb++) sg_load<R, INV>(sr.data(), si.data(), bb, M, Ns, tw, &vr[bb * R], &vi[bb * R]);
    for (int bb = 0; bb < nbf; bb++) sg_store<R>(sr.data(), si.data(), bb, Ns, &vr[bb * R], &vi[bb * R]);
}

template <bool INV>
static void host_fft_g(std::vector<double> &sr, std::vector<double> &si, int M, const double *tw)
{
    int Ns = 1, two = M;
    while (two % 3 == 0) two /= 3;
    for (; two % 8 == 0; two /= 8, Ns *= 8) host_pass_g<8, INV>(sr, si, M, Ns, tw);
    if (two == 4) { host_pass_g<4, INV>(sr, si, M, Ns, tw); Ns *= 4; }
    else if (two == 2) { host_pass_g<2, INV>(sr, si, M, Ns, tw); Ns *= 2; }
    for (; Ns < M; Ns *= 3) host_pass_g<3, INV>(sr, si, M, Ns, tw);
}

template <int LR, bool INV>
static void host_pass_p(std::vector<double> &sr, std::vector<double> &si, int logM, int logNs, const double *tw)
{
    const int R = 1 << LR, nbf = 1 << (logM - LR);
    std::vector<double> vr((size_t)nbf * R), vi((size_t)nbf * R);
    for (int bb = 0; bb < nbf; bb++) sk_load<LR, INV>(sr.data(), si.data(), bb, logM, logNs, tw, &vr[bb * R], &vi[bb * R]);
    for (int bb = 0; bb < nbf; bb++) sk_store<LR>(sr.data(), si.data(), bb, logNs, &vr[bb * R], &vi[bb * R]);
}

template <bool INV>
static void host_fft_p(std::vector<double> &sr, std::vector<double> &si, int logM, const double *tw)
{
    int logNs = 0, rem = logM;
    for (; rem >= 3; rem -= 3, logNs += 3) host_pass_p<3, INV>(sr, si, logM, logNs, tw);
    if (rem == 2) host_pass_p<2, INV>(sr, si, logM, logNs, tw);
    else if (rem == 1) host_pass_p<1, INV>(sr, si, logM, logNs, tw);
}

int main(int argc, char **argv)
{
    const int M = atoi(argv[1]);
    const double PI = 3.14159265358979323846;
    std::vector<double> tw(2 * M), xr(M), xi(M);
    for (int t = 0; t < M; t++) { tw[2 * t] = cos(2 * PI * t / M); tw[2 * t + 1] = -sin(2 * PI * t / M); }
    unsigned long long s = 12345;
    for (int m = 0; m < M; m++) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL; xr[m] = (double)(s >> 11) / 9007199254740992.0 - 0.5;
        s = s * 6364136223846793005ULL + 1442695040888963407ULL; xi[m] = (double)(s >> 11) / 9007199254740992.0 - 0.5;
    }
    int logM = -1;
    for (int l = 0; l < 31; l++) if ((1 << l) == M) logM = l;
    for (int inv = 0; inv < 2; inv++) {
        std::vector<double> ar = xr, ai = xi;
        if (inv) host_fft_g<true>(ar, ai, M, tw.data()); else host_fft_g<false>(ar, ai, M, tw.data());
        std::vector<double> br = xr, bi = xi;
        if (logM >= 0) { if (inv) host_fft_p<true>(br, bi, logM, tw.data()); else host_fft_p<false>(br, bi, logM, tw.data()); }
        double err = 0.0, errp = 0.0, scale = 0.0;
        for (int k = 0; k < M; k++) {
            long double sr_ = 0, si_ = 0;
            for (int m = 0; m < M; m++) {
                long double a = (inv ? 2.0L : -2.0L) * 3.14159265358979323846264338327950288L * (long double)((long long)k * m % M) / M;
                sr_ += xr[m] * cosl(a) - xi[m] * sinl(a);
                si_ += xr[m] * sinl(a) + xi[m] * cosl(a);
            }
            err = fmax(err, fmax(fabs(ar[k] - (double)sr_), fabs(ai[k] - (double)si_)));
            if (logM >= 0) errp = fmax(errp, fmax(fabs(br[k] - (double)sr_), fabs(bi[k] - (double)si_)));
            scale = fmax(scale, fmax(fabs((double)sr_), fabs((double)si_)));
        }
        printf("%s %.3e %.3e\n", inv ? "inverse" : "forward", err / scale, logM >= 0 ? errp / scale : 0.0);
    }
    return 0;
}
