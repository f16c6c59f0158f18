// Symbolic analysis of the mole-reactor preconditioner (see sparse_lu.hpp).
#include "sparse_lu.hpp"

#include <stdexcept>

namespace cg {

SparseLUHost sparse_lu_analyse(int N, const std::vector<char>& pattern) {
    if (N <= 0 || static_cast<int>(pattern.size()) != N * N) throw std::runtime_error("sparse_lu_analyse: bad pattern");
    SparseLUHost s;
    s.N = N;
    // minimum-degree order on the symmetrised elimination graph (ties to the
    // lowest index); eliminating a node joins its remaining neighbours
    std::vector<char> G(static_cast<size_t>(N) * N, 0);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            if (i != j && (pattern[i * N + j] || pattern[j * N + i])) G[i * N + j] = 1;
    std::vector<char> alive(N, 1);
    std::vector<int> nb;
    for (int step = 0; step < N; step++) {
        int best = -1, bestdeg = 0;
        for (int v = 0; v < N; v++) {
            if (!alive[v]) continue;
            int d = 0;
            for (int u = 0; u < N; u++) d += alive[u] && G[v * N + u];
            if (best < 0 || d < bestdeg) { best = v; bestdeg = d; }
        }
        s.perm.push_back(best);
        nb.clear();
        for (int u = 0; u < N; u++)
            if (alive[u] && G[best * N + u]) nb.push_back(u);
        for (int a : nb)
            for (int b : nb)
                if (a != b) G[a * N + b] = 1;
        alive[best] = 0;
    }
    s.iperm.assign(N, 0);
    for (int k = 0; k < N; k++) s.iperm[s.perm[k]] = k;
    // fill pattern in elimination order
    std::vector<char> F(static_cast<size_t>(N) * N, 0);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            if (i == j || pattern[i * N + j]) F[s.iperm[i] * N + s.iperm[j]] = 1;
    for (int k = 0; k < N; k++)
        for (int i = k + 1; i < N; i++)
            if (F[i * N + k])
                for (int j = k + 1; j < N; j++)
                    if (F[k * N + j]) F[i * N + j] = 1;
    // value slots, row-major in elimination order
    std::vector<int> slot(static_cast<size_t>(N) * N, -1);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            if (F[i * N + j]) {
                slot[i * N + j] = s.nnz++;
                s.row.push_back(s.perm[i]);
                s.col.push_back(s.perm[j]);
            }
    s.pos.assign(static_cast<size_t>(N) * N, -1);
    for (int r = 0; r < N; r++)
        for (int c = 0; c < N; c++) s.pos[r * N + c] = slot[s.iperm[r] * N + s.iperm[c]];
    s.diag.resize(N);
    for (int k = 0; k < N; k++) s.diag[k] = slot[k * N + k];
    s.lptr.push_back(0);
    s.optr.push_back(0);
    s.uptr.push_back(0);
    for (int k = 0; k < N; k++) {
        for (int i = k + 1; i < N; i++)
            if (F[i * N + k]) {
                s.lslot.push_back(slot[i * N + k]);
                s.lrow.push_back(i);
                for (int j = k + 1; j < N; j++)
                    if (F[k * N + j]) {
                        s.ops.push_back(slot[i * N + j]);
                        s.ops.push_back(slot[i * N + k]);
                        s.ops.push_back(slot[k * N + j]);
                    }
            }
        for (int i = 0; i < k; i++)
            if (F[i * N + k]) {
                s.uslot.push_back(slot[i * N + k]);
                s.urow.push_back(i);
            }
        s.lptr.push_back(static_cast<int>(s.lslot.size()));
        s.optr.push_back(static_cast<int>(s.ops.size() / 3));
        s.uptr.push_back(static_cast<int>(s.uslot.size()));
    }
    return s;
}

}  // namespace cg
