"""blochparty_amd: an MI355X-native engine for blochparty's ``simulate(..., engine="gpu")`` path.

The inputs (``Phantom``, ``Sequence``, coil sensitivities and the builders that make them) and the output
(``SimResult``) mirror the reference package's; ``simulate`` runs them through hand-written HIP kernels for gfx950,
forward and with gradients (a checkpointed adjoint behind ``torch.autograd.Function``).
"""

from .builder import (
    ADC,
    MAX_RF_STEP,
    RFPulse,
    SequenceBuilder,
    System,
    Trapezoid,
    block_duration,
    block_pulse,
    sinc_pulse,
    trapezoid,
)
from .coils import birdcage_sensitivities
from .exchange import BoundPool
from .fields import GradientNonlinearity, Scanner
from .engines import available, simulate
from .library import gre_2d, gre_3d, gre_cartesian, tse_2d
from .phantom import (
    LABELS,
    TISSUES,
    Phantom,
    grid_positions,
    phantom_from_maps,
    shepp_logan,
    shepp_logan_maps,
    three_class_maps,
)
from .pulseq import read_pulseq
from .recon import coil_combine, reconstruct, rss
from .result import SimResult
from .sequence import Sequence

__all__ = [
    "ADC",
    "BoundPool",
    "GradientNonlinearity",
    "Scanner",
    "LABELS",
    "MAX_RF_STEP",
    "TISSUES",
    "Phantom",
    "RFPulse",
    "Sequence",
    "SequenceBuilder",
    "SimResult",
    "System",
    "Trapezoid",
    "available",
    "birdcage_sensitivities",
    "block_duration",
    "block_pulse",
    "grid_positions",
    "gre_2d",
    "gre_3d",
    "gre_cartesian",
    "phantom_from_maps",
    "read_pulseq",
    "reconstruct",
    "rss",
    "coil_combine",
    "shepp_logan",
    "shepp_logan_maps",
    "simulate",
    "sinc_pulse",
    "three_class_maps",
    "trapezoid",
    "tse_2d",
]
