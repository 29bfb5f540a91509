from .ghost_fields import State, allocate_state, ghost_shapes, interior_view

__all__ = ["State", "allocate_state", "interior_view", "ghost_shapes"]
