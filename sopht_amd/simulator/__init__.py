"""Simulator layer: the 3D flow simulator and the immersed-body glue."""

from .adjoint import (
    CoupledFlowRollout,
    HeaveKinematics,
    IBFlowRolloutFunction,
    body_forces,
    time_averaged_force,
)
from .batch import BatchedFlowSimulation
from .flow import (
    FlowSimulator,
    UnboundedNavierStokesFlowSimulator3D,
    create_unbounded_flow_simulator_3d,
)
from .passive_transport import PassiveTransportFlowSimulator, compute_advection_diffusion_stable_timestep
from .cosserat_rod import (
    CosseratRodElementCentricForcingGrid,
    CosseratRodFlowInteraction,
    CosseratRodNodalForcingGrid,
    CosseratRodSurfaceForcingGrid,
    FlowForces,
    StraightRod,
)
from .immersed_body import (
    FibonacciEllipsoidForcingGrid,
    FibonacciSphereForcingGrid,
    ImmersedBodyFlowInteraction,
    ImmersedBodyForcingGrid,
    RigidBody,
    RigidBodyFlowInteraction,
    SphereForcingGrid,
    ThreeDimensionalRigidBodyForcingGrid,
)

__all__ = [
    "BatchedFlowSimulation",
    "CosseratRodElementCentricForcingGrid",
    "CosseratRodFlowInteraction",
    "CosseratRodNodalForcingGrid",
    "CosseratRodSurfaceForcingGrid",
    "FlowForces",
    "StraightRod",
    "CoupledFlowRollout",
    "HeaveKinematics",
    "IBFlowRolloutFunction",
    "body_forces",
    "time_averaged_force",
    "FibonacciEllipsoidForcingGrid",
    "FibonacciSphereForcingGrid",
    "FlowSimulator",
    "ImmersedBodyFlowInteraction",
    "ImmersedBodyForcingGrid",
    "PassiveTransportFlowSimulator",
    "RigidBody",
    "RigidBodyFlowInteraction",
    "SphereForcingGrid",
    "ThreeDimensionalRigidBodyForcingGrid",
    "UnboundedNavierStokesFlowSimulator3D",
    "compute_advection_diffusion_stable_timestep",
    "create_unbounded_flow_simulator_3d",
]
