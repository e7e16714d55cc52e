"""milc_qcd_amd: MI355X-native HISQ/asqtad staggered Dslash, CG, multi-shift CG
and fermion-force outer products behind a C-ABI that mirrors MILC's generic_ks
GPU interface (include/milc_hip.h).  The compute path is libmilc_hip.so
(hand-written HIP for gfx950); this package only carries the host-side mirror
of the MILC calls used by tests and the benchmark; lattice_io mirrors MILC's
lattice-file readers and writers (include/milc_io.h)."""
from .ks import (EVEN, ODD, EVENANDODD, EXPORTS, KsParam, Lattice, MilcHipError,  # noqa: F401
                 QuarkInvertControl, LIB_PATH, load_library, load_rhmc_params, gauge_loop_coeffs, Rational,
                 Pseudofermion, pseudofermions_from_file, INT_LEAPFROG, INT_OMELYAN, INT_2G1F, INT_3G1F,
                 INT_5G1F, INT_6G1F)

from . import lattice_io  # noqa: F401,E402

# device field precisions of the library (mh_common.h Prec)
P_F64, P_F32, P_H16 = 0, 1, 2
