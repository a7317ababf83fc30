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
