"""What a queue element is, told from the attributes qibo's classes carry.

The backend receives gates built by qibo itself (``qibo.gates.*`` once qibo
loads this package through ``set_backend``) or by this package's own caller
surface, so it never dispatches on class identity. The attributes read are
those of qibo 0.3.5:

* measurement ``M`` (src/qibo/gates/measurements.py:46-60): ``collapse``,
  ``register_name`` and ``result``; its ``matrix()`` raises
  (measurements.py:182-187), so a measurement is never planned as a gate;
* ``FusedGate`` (src/qibo/gates/special.py:34-47): ``gates``, ``qubit_set``
  and ``marked``;
* channels (src/qibo/gates/channels.py:16-22): ``coefficients`` and ``gates``;
* ``CallbackGate`` (src/qibo/gates/special.py:6-25): ``callback``.
"""


def is_measurement(gate) -> bool:
    return hasattr(gate, "collapse") and hasattr(gate, "register_name") and hasattr(gate, "result")


def is_channel(gate) -> bool:
    return getattr(gate, "coefficients", None) is not None and hasattr(gate, "gates")


def is_fused(gate) -> bool:
    return (hasattr(gate, "gates") and hasattr(gate, "qubit_set") and hasattr(gate, "marked")
            and not is_channel(gate))


def is_callback(gate) -> bool:
    return hasattr(gate, "callback")


def is_special(gate) -> bool:
    """A special gate that is not a fused group: it runs on the state through its
    own ``apply`` between the compiled passes (callbacks, and this package's
    other ``SpecialGate`` subclasses)."""
    if is_fused(gate):
        return False
    return is_callback(gate) or isinstance(gate, _special_base())


_SPECIAL = []


def _special_base():
    if not _SPECIAL:  # imported on first use: gates imports this module
        from .gates import SpecialGate

        _SPECIAL.append(SpecialGate)
    return _SPECIAL[0]


def runs_outside_passes(gate) -> bool:
    """Collapsing measurements, channels and special gates: applied one by one
    through ``gate.apply`` between blocked passes."""
    if is_measurement(gate):
        return bool(gate.collapse)
    return is_channel(gate) or is_special(gate)
