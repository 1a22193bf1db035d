// Host-side check of the XCD block remaps (ofx_common.h): for random launch
// shapes and segment cuts, xcd_remap_seg must be a bijection on [0, nblk)
// whose blocks of segment s stay inside s, and with one segment it must equal
// xcd_remap; for power-of-two block sizes the shift form xcd_remap_seg_lg
// must agree with it block for block.  Prints "ok <blocks>" or the first failure; runs no GPU code.
#include "../oneflow_amd/csrc/ofx_common.h"

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

using namespace ofx;

int main() {
    std::mt19937 rng(7);
    long checked = 0;
    for (int trial = 0; trial < 20000; ++trial) {
        int per = (trial % 3 == 0) ? 64 : (trial % 3 == 1 ? 256 : 37);
        int n = 1 + (int)(rng() % 50000);
        int nseg = 1 + (int)(rng() % MAXSEG);
        std::vector<int> cut;
        for (int s = 1; s < nseg; ++s) cut.push_back((int)(rng() % (n + 1)));
        std::sort(cut.begin(), cut.end());
        std::vector<int> st(MAXSEG + 1, 0);
        for (int s = 1; s < nseg; ++s) st[s] = cut[s - 1];
        st[nseg] = n;
        int nblk = (n + per - 1) / per;
        std::vector<int> seen(nblk, 0);
        for (int b = 0; b < nblk; ++b) {
            int lb = xcd_remap_seg(b, nblk, per, st.data(), nseg);
            if (lb < 0 || lb >= nblk) { std::printf("out of range: trial %d block %d -> %d\n", trial, b, lb); return 1; }
            ++seen[lb];
            if (nseg == 1 && lb != xcd_remap(b, nblk)) { std::printf("one segment differs: trial %d\n", trial); return 1; }
            if ((per & (per - 1)) == 0 && lb != xcd_remap_seg_lg(b, nblk, __builtin_ctz((unsigned)per), st.data(), nseg)) {
                std::printf("shift form differs: trial %d block %d\n", trial, b);
                return 1;
            }
        }
        for (int b = 0; b < nblk; ++b)
            if (seen[b] != 1) {
                std::printf("not a bijection: trial %d n %d per %d nseg %d\n", trial, n, per, nseg);
                return 1;
            }
        checked += nblk;
    }
    std::printf("ok %ld\n", checked);
    return 0;
}
