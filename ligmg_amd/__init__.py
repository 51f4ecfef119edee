"""MI355X-native graph-Laplacian algebraic multigrid (LigMG setup + solve)."""
from . import graphs  # noqa: F401
