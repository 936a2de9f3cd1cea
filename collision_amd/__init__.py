"""collision_amd: the collision step of a cloth / string / rigid-body solver
(proximity, continuous collision detection, impulse response, impact zones,
strain limiting) as hand-written HIP kernels for MI355X (gfx950), behind a
C ABI (include/collision_amd.h) that mirrors the reference CollisionSolver3d
interface."""
from . import scenes  # noqa: F401

__all__ = ["scenes", "CollisionSolver3d", "solver_for_scene", "CollisionError"]


def __getattr__(name):
    if name in ("CollisionSolver3d", "solver_for_scene", "CollisionError", "default_params"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
