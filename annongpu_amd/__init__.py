"""annongpu_amd -- MI355X-native variational Monte Carlo path of ANNonGPU.

Drop-in for the reference's pyANNonGPU on the VMC path
(``import annongpu_amd as pyANNonGPU``): wave functions PsiRBM / PsiDeep /
PsiCNN / PsiClassical, Pauli-sum Operators, MonteCarlo and ExactSummation
ensembles over spins and Pauli strings, ExpectationValue, TDVP (+ the SR solve),
HilbertSpaceDistance, KullbackLeibler, a learned block proposal
(BlockProposal) and the free functions of
pyANNonGPU/main.cpp.template:494-540, all computed by hand-written gfx950 HIP
kernels behind the C ABI in include/annongpu.h.
"""
from .pauli import PauliExpression, sigma_x, sigma_y, sigma_z, identity
from .api import (
    Spins, PauliString, PsiRBM, PsiDeep, PsiCNN, PsiClassical, PsiFullyPolarized, PsiClassicalFP_1,
    PsiClassicalFP_2, PsiClassicalANN_1, PsiClassicalANN_2, new_classical_network, Operator, MonteCarloSpins,
    ExactSummationSpins, MonteCarloPaulis, ExactSummationPaulis, SampleList, PauliSampleList,
    ExpectationValue, TDVP, HilbertSpaceDistance, KullbackLeibler, log_psi_s, psi_O_k, psi_O_k_vector, log_psi, psi_vector,
    log_psi_vector, apply_operator, local_energies, activation_function, setDevice,
    start_profiling, stop_profiling, profiler_query, profiler_reset, BlockProposal, rect_blocks, chain_blocks,
)
from .constructors import (
    new_RBM, new_deep_neural_network, new_convolutional_network, seeded_rbm, seeded_rbm_deep,
    seeded_cnn, deep_connections,
)
from .json_numpy import NumpyEncoder, NumpyDecoder
from . import hamiltonians
from .evolution import RK4, Euler

__all__ = [
    "PauliExpression", "sigma_x", "sigma_y", "sigma_z", "identity", "Spins", "PauliString", "PsiRBM", "PsiDeep",
    "PsiCNN", "PsiClassical", "PsiFullyPolarized", "PsiClassicalFP_1", "PsiClassicalFP_2", "PsiClassicalANN_1",
    "PsiClassicalANN_2", "new_classical_network", "Operator", "MonteCarloSpins", "ExactSummationSpins",
    "MonteCarloPaulis", "ExactSummationPaulis", "SampleList", "PauliSampleList", "ExpectationValue", "TDVP",
    "HilbertSpaceDistance", "KullbackLeibler", "log_psi_s", "psi_O_k",
    "psi_O_k_vector", "log_psi", "psi_vector", "log_psi_vector", "apply_operator", "local_energies",
    "activation_function", "setDevice", "start_profiling", "stop_profiling", "profiler_query", "profiler_reset",
    "new_RBM", "new_deep_neural_network", "new_convolutional_network", "seeded_rbm", "seeded_rbm_deep",
    "seeded_cnn", "deep_connections", "NumpyEncoder", "NumpyDecoder", "hamiltonians", "BlockProposal",
    "rect_blocks", "chain_blocks", "RK4", "Euler",
]
