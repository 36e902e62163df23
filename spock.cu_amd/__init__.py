"""MI355X-native SPOCK: offline generator (``api``) and the HIP solver runtime (``runtime``)."""
