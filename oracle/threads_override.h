// Forced include (g++ -include) of the all-cores reference build: the
// reference's configuration header, then its OpenMP thread count raised
// from the fixed 4 (config/config.h:28-29, used by every num_threads clause)
// to PPREF_THREADS.  The header's include guard keeps the sources' own
// inclusion of it from restoring the old value.  Test infrastructure only:
// the CPU baseline of bench.py times this build beside the 4-thread one.
#include "config/config.h"
#undef THREAD_NUM
#define THREAD_NUM PPREF_THREADS
#undef MAX_THREAD_NUM
#define MAX_THREAD_NUM PPREF_THREADS
