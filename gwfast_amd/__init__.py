"""gwfast_amd: MI355X-native event-batched waveform / SNR / Fisher path with
gwfast's public API (waveforms, signal, network, fisherTools)."""
