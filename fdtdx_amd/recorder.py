"""Interface recorder (reference: src/fdtdx/interfaces/recorder.py:14-199).

Only the exact (``modules=[]``) recording is implemented: the engine stores every step's PML
interface layer in float32 on the device.
"""

from dataclasses import dataclass, field


@dataclass
class Recorder:
    modules: list = field(default_factory=list)

    def __post_init__(self):
        if self.modules:
            raise NotImplementedError("recording compression modules are not supported by this engine")
