// Host-side HSS planning: tree of one merge (aux_hss.f90:737-827 npart/n2tree/
// btree/child, GetTreeInfo :238-307) and the level tables the kernels consume.
#pragma once
#include <math.h>
#include <vector>

struct HssTree {
    int K = 0, ni = 0, ltr = 0, lm = 0, lvl = 0;
    std::vector<int> tr;      // parent (1-based, 0 = root), size ltr
    std::vector<int> m;       // leaf sizes (lm)
    std::vector<int> c1, c2;  // children (1-based node ids, 0 = none)
    std::vector<int> depth;   // root depth 0
    std::vector<std::vector<int>> lev;  // lev[LL] = nodes at level LL (LL = 1..lvl), ascending id
    std::vector<int> leaf_r0; // per node: leaf first index - 1
};

static void hss_btree(int n, int *tr)
{
    if (n > 3) {
        int m = (n - 1) / 2;
        hss_btree(m, tr);
        for (int i = 1; i <= m; i++) tr[m + i - 1] = tr[i - 1] + m;
        tr[m - 1] = n;
        tr[2 * m - 1] = n;
        tr[n - 1] = 0;
    } else {
        tr[0] = 3; tr[1] = 3; tr[2] = 0;
    }
}

static void hss_n2tree(int n, int *tr)
{
    if (n == 1) { tr[0] = 0; return; }
    int n1 = (1 << (int)floorf(logf((float)n) / logf(2.0f))) - 1;
    hss_btree(n1, tr);
    hss_n2tree(n - n1 - 1, tr + n1);
    for (int i = n1 + 1; i <= n - 1; i++) tr[i - 1] += n1;
    tr[n1 - 1] = n;
    tr[n - 2] = n;
    tr[n - 1] = 0;
}

// returns false when the tree is a single leaf (no off-diagonal blocks)
static bool hss_make_tree(HssTree &T, int K, int ni)
{
    T.K = K; T.ni = ni;
    int l = K / ni;
    if (l < 1) return false;
    T.m.assign(l + 1, ni);
    int mr = K % ni;
    if (mr >= ni / 2) { l++; T.m[l - 1] = mr; }
    else T.m[l - 1] += mr;
    T.m.resize(l);
    T.lm = l;
    T.ltr = 2 * l - 1;
    if (T.ltr < 3 || l > 512) return false;  // compaction tables hold <= 1024 nodes
    T.tr.assign(T.ltr + 8, 0);
    hss_n2tree(T.ltr, T.tr.data());
    T.tr.resize(T.ltr);
    T.c1.assign(T.ltr + 1, 0);
    T.c2.assign(T.ltr + 1, 0);
    for (int i = 1; i <= T.ltr - 1; i++) {
        int p = T.tr[i - 1];
        if (T.c1[p] == 0) T.c1[p] = i; else T.c2[p] = i;
    }
    T.depth.assign(T.ltr + 1, 0);
    int lvl = 0;
    for (int i = T.ltr - 1; i >= 1; i--) {
        T.depth[i] = T.depth[T.tr[i - 1]] + 1;
        lvl = std::max(lvl, T.depth[i]);
    }
    T.lvl = lvl;
    T.lev.assign(lvl + 1, std::vector<int>());
    for (int j = 1; j <= T.ltr - 1; j++) T.lev[T.depth[j]].push_back(j);
    T.leaf_r0.assign(T.ltr + 1, 0);
    int st = 0, it = 0;
    for (int i = 1; i <= T.ltr; i++)
        if (T.c1[i] == 0 && T.c2[i] == 0) { T.leaf_r0[i] = st; st += T.m[it++]; }
    return true;
}
