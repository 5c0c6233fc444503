// bk_sweeps.hpp -- which qubits each HIGH sweep of a Trotter step rotates.
#pragma once

#include <algorithm>
#include <numeric>
#include <vector>

namespace bk {

struct HighSweep {
    int h1, a1, h2, a2, R;
};

// Split the qubits [tb, n) into HIGH sweeps of at most htb - rmin qubits.
// page_bits > 0: qubits >= page_bits move rows across pages (2^page_bits
// amplitudes per page); they are dealt out evenly over the sweeps, each sweep
// taking a contiguous run of the low qubits plus a contiguous run of the
// page-crossing ones (two segments).  page_bits == 0: contiguous runs.
inline std::vector<HighSweep> plan_high_sweeps(int n, int tb, int htb, int rmin, int page_bits)
{
    std::vector<HighSweep> out;
    const int remaining = n - tb;
    if (remaining <= 0) return out;
    const int amax = htb - rmin;
    const int nh = (remaining + amax - 1) / amax;
    std::vector<int> A(nh);
    for (int i = 0; i < nh; ++i) A[i] = (remaining - std::accumulate(A.begin(), A.begin() + i, 0) + (nh - i) - 1) / (nh - i);
    const int pb = (page_bits > tb && page_bits < n) ? page_bits : n;  // first page-crossing qubit
    const int nlow = pb - tb, nhigh = n - pb;
    if (page_bits <= 0 || nh == 1 || nhigh <= 0) {
        int h = tb;
        for (int i = 0; i < nh; ++i) {
            out.push_back({h, A[i], h + A[i], 0, htb - A[i]});
            h += A[i];
        }
        return out;
    }
    // page-crossing qubits per sweep (even split, larger first), the rest low
    std::vector<int> hi(nh);
    for (int i = 0; i < nh; ++i) hi[i] = (nhigh - std::accumulate(hi.begin(), hi.begin() + i, 0) + (nh - i) - 1) / (nh - i);
    int lo_pos = tb, hi_pos = n;
    for (int i = 0; i < nh; ++i) {
        int hcnt = std::min(hi[i], A[i]);
        int lcnt = A[i] - hcnt;
        if (lo_pos + lcnt > pb) {  // out of low qubits: take more high ones
            lcnt = pb - lo_pos;
            hcnt = A[i] - lcnt;
        }
        hi_pos -= hcnt;
        if (lcnt > 0 && hcnt > 0)
            out.push_back({lo_pos, lcnt, hi_pos, hcnt, htb - A[i]});
        else if (lcnt > 0)
            out.push_back({lo_pos, lcnt, lo_pos + lcnt, 0, htb - A[i]});
        else
            out.push_back({hi_pos, hcnt, hi_pos + hcnt, 0, htb - A[i]});
        lo_pos += lcnt;
    }
    (void)nlow;
    return out;
}

}  // namespace bk
