// ise_format.cpp -- the reference's text output of enumeration results:
// printed matrices (CMatrixData::printOut, Source/Matrix.cpp:43-155) and the
// |Aut(D)| summary table (CGroupsInfo::printGroupInfo, Source/GroupsInfo.cpp:
// 96-150, with CNumbInfo::outNumbInfo, :9-16), so that result files can be
// compared with the reference's line by line.
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <array>
#include <map>
#include <string>
#include <vector>
#include "ise.h"

namespace {

struct Out {
    char *buf;
    uint64_t cap, len = 0;
    void put(const std::string &s) {
        if (buf && len + s.size() < cap) memcpy(buf + len, s.data(), s.size());
        len += s.size();
    }
    int finish(uint64_t *len_out) {
        if (len_out) *len_out = len;
        if (!buf || len >= cap) return ISE_ERR_CAPACITY;
        buf[len] = 0;
        return ISE_OK;
    }
};

std::string fmt(const char *f, unsigned long long x) {
    char t[64];
    snprintf(t, sizeof t, f, x);
    return t;
}

} // namespace

extern "C" {

int ise_format_matrix(const uint64_t *rows, int nrows, int ncols, uint64_t number, uint64_t sub, int canon,
                      uint64_t aut, char *buf, uint64_t cap, uint64_t *len_out) {
    if (!rows || nrows < 1 || nrows > ISE_MAX_ROWS || ncols < 1 || ncols > ISE_MAX_COLS) return ISE_ERR_ARG;
    Out o{buf, cap};
    if (number > 0) {
        std::string h = fmt("\nMatrix # %3llu", number);
        if (sub) h += fmt(".%llu", sub);
        if (!canon) h += "  *** Non-canonical ***";
        else if (canon < 0) h += "  *** Canonicity was not tested ***";
        h += fmt("    |Aut(M)| = %6llu", aut);
        o.put(h + "\n");
    }
    std::string line(ncols, '0');
    for (int i = 0; i < nrows; i++) {
        for (int c = 0; c < ncols; c++) line[c] = ((rows[i] >> c) & 1) ? '1' : '0';
        o.put(line + "\n");
    }
    return o.finish(len_out);
}

int ise_format_aut_table(int v, int b, int r, int k, int lambda, uint32_t n, const uint64_t *designs,
                         const uint64_t *auts, const uint8_t *orbits, char *buf, uint64_t cap, uint64_t *len_out) {
    if (v < 1 || v > ISE_MAX_ROWS || b < 1 || b > ISE_MAX_COLS || (n && (!designs || !auts || !orbits)))
        return ISE_ERR_ARG;
    // per |Aut|: designs, simple (no replicated blocks), transitive on the rows, both
    std::map<uint64_t, std::array<unsigned long long, 4>> rowsByOrder;
    std::array<unsigned long long, 4> total{0, 0, 0, 0};
    for (uint32_t d = 0; d < n; d++) {
        const uint64_t *m = designs + (size_t)d * v;
        std::vector<uint32_t> cols(b, 0);
        for (int i = 0; i < v; i++)
            for (int c = 0; c < b; c++) cols[c] |= (uint32_t)((m[i] >> c) & 1) << i;
        std::sort(cols.begin(), cols.end());
        const bool simple = std::adjacent_find(cols.begin(), cols.end()) == cols.end();
        bool transitive = true;
        for (int i = 0; i < v; i++) transitive &= orbits[(size_t)d * ISE_MAX_ROWS + i] == 0;
        auto &row = rowsByOrder[auts[d]];
        const unsigned long long add[4] = {1, simple, transitive, simple && transitive};
        for (int t = 0; t < 4; t++) {
            row[t] += add[t];
            total[t] += add[t];
        }
    }
    Out o{buf, cap};
    char head[64];
    snprintf(head, sizeof head, "BIBD(%3d, %3d, %2d, %2d, %2d)\n", v, b, r, k, lambda);
    o.put(head);
    const std::string title = "\n        |Aut(D)|          Nd:           Ns:   Ndt:   Nst:\n";
    o.put(title);
    const std::string rule = "    " + std::string(title.size(), '_') + "\n";
    o.put(rule);
    auto counts = [](const std::array<unsigned long long, 4> &c) {
        return fmt("   %12llu", c[0]) + fmt("   %10llu", c[1]) + fmt("   %4llu", c[2]) + fmt("   %4llu", c[3]);
    };
    for (auto &kv : rowsByOrder) o.put(fmt("    %10llu", kv.first) + counts(kv.second) + "\n");
    o.put(rule);
    o.put("        Total:" + counts(total) + "\n");
    return o.finish(len_out);
}

} // extern "C"

// ---- 1-factorizations as the k-Sys result files print them ----------------
// (TableAut::name + Table::printTable + outMatrix,
// k-Factorization/TripleSys/Include/Table.h:5-41, 93-115, 333-361, with the
// cycle statistics of alldata::matrixStatOutput / collectCyclesAndPath,
// TripleSys/Source/p1fCheck.cpp:77-118, 499-517).  Per matrix:
//   "\n\n%5d: |Aut(M)| =  <aut>,  Cycles:<sets>\n"
// where every pair of rounds contributes its cycle type (the lengths of the
// cycles of the union of the two perfect matchings, ascending) and the sets are
// "<count>(<l1>:<l2>...)" in increasing order of their length lists, then one
// line per round, ' "' + for every pair " %3d%3d" + ' "', then an empty line.
extern "C" int ise_kf_format_matrix(const uint64_t *rounds, int nrounds, int players, uint64_t number, uint64_t aut,
                                    char *buf, uint64_t cap, uint64_t *len_out) {
    if (!rounds || players < 4 || players > 16 || (players & 1) || nrounds < 1 || nrounds > players - 1)
        return ISE_ERR_ARG;
    const int n = players;
    std::vector<std::array<uint8_t, 16>> part(nrounds);
    for (int r = 0; r < nrounds; r++) {
        uint32_t seen = 0;
        for (int p = 0; p < n; p += 2) {
            const int x = (int)(rounds[r] >> (4 * p)) & 0xf, y = (int)(rounds[r] >> (4 * p + 4)) & 0xf;
            if (x >= n || y >= n || x == y || (seen >> x & 1) || (seen >> y & 1)) return ISE_ERR_ARG;
            seen |= (1u << x) | (1u << y);
            part[r][x] = (uint8_t)y;
            part[r][y] = (uint8_t)x;
        }
    }
    // cycle type of every pair of rounds, counted per type
    std::map<std::vector<int>, uint64_t> types;
    for (int m = 1; m < nrounds; m++)
        for (int i = 0; i < m; i++) {
            uint32_t used = 0;
            std::vector<int> lens;
            for (int s = 0; s < n; s++) {
                if (used >> s & 1) continue;
                int v = s, len = 0;
                do {
                    const int w = part[i][v];
                    used |= (1u << v) | (1u << w);
                    len += 2;
                    v = part[m][w];
                } while (v != s && len <= n);
                lens.push_back(len);
            }
            std::sort(lens.begin(), lens.end());
            types[lens]++;
        }
    std::string head = fmt("%5llu: |Aut(M)| = ", number) + fmt(" %llu", aut) + ",  Cycles:";
    bool first = true;
    for (const auto &t : types) {
        if (!first) head += " ";
        first = false;
        head += fmt("%llu(", t.second);
        for (size_t j = 0; j < t.first.size(); j++) head += (j ? ":" : "") + fmt("%llu", (unsigned long long)t.first[j]);
        head += ")";
    }
    Out o{buf, cap};
    o.put("\n\n" + head + "\n");
    for (int r = 0; r < nrounds; r++) {
        std::string line = " \"";
        for (int p = 0; p < n; p++) {
            if (!(p & 1)) line += " ";
            line += fmt(" %2llu", (rounds[r] >> (4 * p)) & 0xf);
        }
        o.put(line + " \"\n");
    }
    o.put("\n");
    return o.finish(len_out);
}
