// cd_tools.cpp -- createCanonizer / releaseCanonizer / runCanonizer
// (k-Factorization/Utils/Include/CDTools.h:9-11, Source/CDTools.cpp:5-51) on
// the library's ise_cd_* entry points.
#include "cd_tools.h"

#include "ise.h"

void *createCanonizer(int v, int lenGroup) {
    if (v <= 0 || lenGroup <= 0) return nullptr;
    ise_cd_canonizer *h = nullptr;
    if (ise_cd_create_canonizer(&h, 0, (uint32_t)v, (uint32_t)lenGroup) != ISE_OK) return nullptr;
    return h;
}

void releaseCanonizer(void *pCanonizer) {
    if (pCanonizer) ise_cd_release_canonizer(static_cast<ise_cd_canonizer *>(pCanonizer));
}

const unsigned char *runCanonizer(void *pCanonizer, const unsigned char *pMatrix, int k, int iDay) {
    if (!pCanonizer || !pMatrix || k <= 0) return nullptr;
    const uint8_t *out = nullptr;
    uint32_t ncols = 0;
    if (ise_cd_run_canonizer(static_cast<ise_cd_canonizer *>(pCanonizer), pMatrix, (uint32_t)k, iDay, &out,
                             &ncols) != ISE_OK)
        return nullptr;
    return out;
}
