// The table-addressed executes of the request server (fmmfft::execute_table:
// a batch of separate caller buffers whose addresses the kernels read from a
// device table), compiled apart from fmmfft.hip so that the two sets of
// kernel instances build in parallel.
#define FMMFFT_TAB_UNIT 1
#include "fmmfft.hip"
