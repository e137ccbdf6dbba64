"""MI355X-native constant-density acoustic time solver with PySIT's API.

The public names mirror PySIT's for this path (``RectangularDomain``,
``PML``, ``CartesianMesh``, ``PointSource``, ``PointReceiver``,
``ReceiverSet``, ``RickerWavelet``, ``Shot``, ``equispaced_acquisition``,
``ConstantDensityAcousticWave``, ``TemporalModeling``,
``generate_seismic_data``, ``TemporalLeastSquares``, ``LBFGS``), so scripts
written against PySIT switch by changing the import.  The time stepping runs
in the HIP library ``libpysit_amd.so`` (gfx950); host code only sets up
geometry and drives it.
"""
from .core import *  # noqa: F401,F403
from .solvers import *  # noqa: F401,F403
from .modeling import *  # noqa: F401,F403
from .objective_functions import *  # noqa: F401,F403
from .optimization import *  # noqa: F401,F403

__version__ = "0.1.0"
