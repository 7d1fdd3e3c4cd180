"""Drop-in alias: `import ghostlight` resolves to the MI355X build."""
from ghostlight_amd import *  # noqa: F401,F403
from ghostlight_amd import _cuda_available  # noqa: F401
