// cd_tools.h -- the k-Sys binary canonizer under the names and signatures of
// k-Factorization/Utils/Include/CDTools.h:9-11 (C++ linkage, as there; the
// reference's LIBRARY_API export macro is a Windows DLL attribute and has no
// counterpart here).  Implemented in cd_tools.cpp on the library's ise_cd_*
// entry points; link that file and libise_amd.so in place of the reference's
// CDTools library.
#pragma once

// createCanonizer(v, k): a canonizer for k-systems of v players in groups of
// lenGroup on HIP device 0; NULL when v, lenGroup are not a k-system's or no
// device is usable.
void *createCanonizer(int v, int lenGroup);

// releaseCanonizer: frees what createCanonizer made (NULL is ignored).
void releaseCanonizer(void *pCanonizer);

// runCanonizer: the canonical binary matrix (v rows of b = days * v / k bytes)
// of the first iDay days of pMatrix (iDay <= 0: all of them); the pointer
// stays valid until the next call on the canonizer.  NULL on invalid input.
const unsigned char *runCanonizer(void *pCanonizer, const unsigned char *pMatrix, int k, int iDay = 0);
