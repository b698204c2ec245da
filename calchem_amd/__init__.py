"""calchem_amd — MI355X-native DFT exchange-correlation quadrature for CalChem."""
