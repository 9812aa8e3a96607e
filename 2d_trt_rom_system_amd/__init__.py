"""MI355X-native full-order TRT iteration (sweep, QD factors, MLOQD/GLOQD, material update)."""
from . import problem  # noqa: F401
from . import native  # noqa: F401
from . import serve  # noqa: F401
