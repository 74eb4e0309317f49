"""Learned denoiser for low-quality renders (build-owned; no reference
counterpart): a small U-Net trained on (low-quality, quality-4000) pairs of
synthetic flames, run on the GPU in BF16 with MFMA kernels."""
