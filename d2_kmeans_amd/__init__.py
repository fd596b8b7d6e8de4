"""MI355X engine for the AD2-clustering hot path (B-ADMM barycenter update,
pdist2 ground distances, Elkan-bounded assignment with exact transport
distances).  The compute lives in ``libd2mi.so`` (HIP kernels for gfx950
behind a C ABI, ``include/d2mi.h``); this package is its Python host side."""
from .data import (BADMM_CEN_OPTIONS, BADMM_CLU_OPTIONS, BADMMOptions, Centroids, D2Data, Phase,
                   D2_EUCLIDEAN_L2, D2_HISTOGRAM, D2_SPARSE_HISTOGRAM, D2_WORD_EMBED,
                   read_d2, write_d2, write_labels)
