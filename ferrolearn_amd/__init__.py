"""ferrolearn_amd -- MI355X (gfx950) squared-Euclidean path of ferrolearn.

KMeans Lloyd (assignment, inertia, centroid update) and brute-force
KNeighborsClassifier / KNeighborsRegressor / NearestNeighbors queries on hand-written HIP
kernels behind the C ABI in include/ferrolearn_hip.h.  Labels, iteration
counts and neighbour indices match the reference's f64 CPU path exactly.
"""
from . import device
from .bisecting import BisectingKMeans, BisectingStrategy, FittedBisectingKMeans
from .cluster import FittedKMeans, KMeans, MiniBatchKMeans, kmeans_fit_batched, kmeans_plusplus
from .device import DeviceArray, to_device
from .errors import (DeviceError, FerroError, InsufficientSamples, InvalidParameter,
                     NumericalInstability, ShapeMismatch)
from .neighbors import (Algorithm, CsrMatrix, FittedKNeighborsClassifier,
                        FittedKNeighborsRegressor, FittedNearestNeighbors, GraphMode,
                        KNeighborsClassifier, KNeighborsRegressor, NearestNeighbors, Weights,
                        kneighbors_graph)
from .metrics import (Metric, euclidean_distances, pairwise_distances_argmin,
                      pairwise_distances_argmin_min)
from .router import CellRouter
from .spectral import FittedSpectralClustering, SpectralClustering

__all__ = [
    "KMeans", "MiniBatchKMeans", "FittedKMeans", "kmeans_plusplus", "KNeighborsClassifier",
    "KNeighborsRegressor", "FittedKNeighborsClassifier", "FittedKNeighborsRegressor",
    "Algorithm", "Weights", "FerroError", "ShapeMismatch", "InsufficientSamples",
    "InvalidParameter", "NumericalInstability", "DeviceError", "DeviceArray", "to_device",
    "device", "CellRouter", "kmeans_fit_batched", "NearestNeighbors", "FittedNearestNeighbors",
    "GraphMode", "CsrMatrix", "kneighbors_graph", "Metric", "euclidean_distances",
    "pairwise_distances_argmin",
    "pairwise_distances_argmin_min", "SpectralClustering", "FittedSpectralClustering",
    "BisectingKMeans", "BisectingStrategy", "FittedBisectingKMeans",
]
