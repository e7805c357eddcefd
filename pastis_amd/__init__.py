"""MI355X batched pairwise protein aligner (HIP, gfx950).

The aligner replaces PASTIS's PWAlign implementations (SeqanFullAligner for
--sfa, SeqanXdropAligner for --sxa) and the pw_aln_batch loop behind the same
operator signature; see include/pastis_hip.h for the C ABI.
"""
from . import _lib  # noqa: F401
