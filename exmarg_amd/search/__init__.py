"""
The analysis of segments with the stored tables on an MI355X GPU: the drop-in for
exmarg.search with device='hip' (src/exmarg/search/__init__.py:1-18).

    store = load_redshift_store(store_dir)                  # once: the tables of every signal power node
    bank = mass_bank(bank_3col_file, segments, 'hip')       # once per frequency grid and noise curve
    res = pe_intrinsic(segments, bank, store, prior)        # M, q, chi, u, ln_lambda, ln_b

    bank = mass_bank(bank_file, segments, 'hip')            # equal mass, zero spin
    res = pe_redshift(segments, bank, store)                # m1, z, ln_lambda, ln_b
"""

from .segments import D0, read_segments
from .bank import mass_bank, MassBank
from .store import load_redshift_store, synthetic_store
from .pe import pe_redshift, pe_intrinsic, distance_weights
from .batch import pe_intrinsic_batch, SegmentBatch
from .serve import Server
from .stream import pe_redshift_stream, psd_estimate, noise_periodograms
