// chkn_digest.cpp -- the digest of the sources this library was built from.
//
// The Makefile hashes csrc/*.h, *.hip and *.cpp (sorted by name) and include/chkn.h and
// rebuilds this file whenever one of them changes, so a library that does not match the
// tree it sits in can be told apart (a stale build copied beside newer sources).
#include "../../include/chkn.h"

#ifndef CHKN_SRC_DIGEST
#error "CHKN_SRC_DIGEST is defined by the Makefile"
#endif

extern "C" const char* chkn_source_digest(void) { return CHKN_SRC_DIGEST; }
