"""Exception taxonomy of the reference (util/exceptions.py:6-126)."""


class UnsupportedRequest(Exception):
    """Raised for requests the engine does not support (util/exceptions.py:6-40)."""

    def __init__(self, request, function, override=False):
        super().__init__(request, function)
        self.request = request
        self.function = function
        self.override = override

    def __str__(self):
        if self.override:
            return "{} in the {} function.".format(self.request, self.function)
        return "The current code does not support {} in the {} function.".format(
            self.request, self.function)


class LockedException(Exception):
    """Raised when a locked object is modified (util/exceptions.py:43-75).

    The reference raises it with a single argument from
    HopsTrajectory.initialize (trajectory/hops_trajectory.py:335-336); here the
    class name is optional so that call site produces this exception."""

    def __init__(self, func_name, class_name=None):
        super().__init__(func_name)
        self.func_name = func_name
        self.class_name = class_name

    def __str__(self):
        return "The {} is currently locked - the attempt to access {} has been canceled.".format(
            self.class_name or "object", self.func_name)


class AuxError(Exception):
    """Misconfigured auxiliary vector (util/exceptions.py:78-101)."""

    def __init__(self, sub_msg):
        super().__init__(sub_msg)
        self.sub_msg = sub_msg

    def __str__(self):
        return "There is a problem in the hierarchy: " + self.sub_msg


class TrajectoryError(Exception):
    """Misconfigured trajectory (util/exceptions.py:104-126)."""

    def __init__(self, sub_msg):
        super().__init__(sub_msg)
        self.sub_msg = sub_msg

    def __str__(self):
        return "There is a problem in the trajectory: " + self.sub_msg


class EngineError(RuntimeError):
    """A call into the HIP engine failed; the message is the engine's own."""


class ArenaFull(EngineError):
    """A ragged-batch step could not place the new bases of some slots in the
    device arena; those slots are marked out of memory and left mid-step,
    every other slot completed the step."""
