"""Drop-in alias: `import ghostlight` resolves to the MI355X build."""
import ghostlight_amd as _ghostlight  # noqa: F401  (the reference's compiled-module name)
from ghostlight_amd import *  # noqa: F401,F403
from ghostlight_amd import _cuda_available, _render_starburst_debug, _render_veil_debug  # noqa: F401
from ghostlight_amd import _solve_source_map, _source_map_landing, _source_map_screen  # noqa: F401
from ghostlight_amd import _aperture_edge_distance_debug, _aperture_sdf_debug, _hurb_sample_debug  # noqa: F401
from ghostlight_amd import _gate_params_debug, _gate_sample_debug, _render_gate_debug  # noqa: F401
