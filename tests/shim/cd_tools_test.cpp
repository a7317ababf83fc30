// Drives the reference-named k-Sys canonizer (shim/cd_tools.h) from text:
// each case is "v k iDay days" followed by days x v player numbers; the
// output per case is the v x b canonical matrix, one line of 0/1 per row,
// then a blank line ("null" when runCanonizer returns NULL).  A canonizer is
// created per (v, k) and kept across the cases that share it.
#include <cstdio>
#include <map>
#include <utility>
#include <vector>

#include "cd_tools.h"

int main() {
    std::map<std::pair<int, int>, void *> canonizers;
    int v, k, iday, days;
    while (std::scanf("%d %d %d %d", &v, &k, &iday, &days) == 4) {
        std::vector<unsigned char> m((size_t)days * v);
        for (auto &x : m) {
            int p;
            if (std::scanf("%d", &p) != 1) return 2;
            x = (unsigned char)p;
        }
        void *&h = canonizers[{v, k}];
        if (!h) h = createCanonizer(v, k);
        if (!h) {
            std::printf("null\n\n");
            continue;
        }
        const unsigned char *out = runCanonizer(h, m.data(), k, iday);
        if (!out) {
            std::printf("null\n\n");
            continue;
        }
        const int b = days * (v / k);
        for (int i = 0; i < v; i++) {
            for (int q = 0; q < b; q++) std::putchar(out[i * b + q] ? '1' : '0');
            std::putchar('\n');
        }
        std::putchar('\n');
    }
    for (auto &kv : canonizers) releaseCanonizer(kv.second);
    return 0;
}
