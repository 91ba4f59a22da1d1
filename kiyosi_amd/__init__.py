"""kiyosi_amd: a Monte Carlo pricing engine for one AMD Instinct MI355X (gfx950).

The engines below take the same contracts, ``PricingContext`` and pricing
calls as kiyosi's CPU Monte Carlo engines and run them as hand-written HIP
kernels through the C ABI in include/kiyosi_amd.h.
"""

from .engines import (
    Book,
    MonteCarloAccumulatorEngine,
    MonteCarloAsianEngine,
    MonteCarloAutocallableEngine,
    MonteCarloBarrierEngine,
    MonteCarloDigitalEngine,
    MonteCarloEstimate,
    MonteCarloVanillaEngine,
)
from .errors import (
    ComputationFailureError,
    DeviceError,
    InvalidInputError,
    KiyosiError,
    MissingHistoryError,
    NonConvergenceError,
    RootNotBracketedError,
    UnsupportedOperationError,
)
from .instruments import (
    Accumulator,
    AmericanOption,
    ArithmeticAveragePriceOption,
    BarrierOption,
    BinarySnowballOption,
    GeometricAveragePriceOption,
    CouponObservation,
    DigitalOption,
    DoubleBarrierOption,
    EuropeanOption,
    KnockInObservation,
    KnockOutObservation,
    PhoenixOption,
    SnowballObservation,
    SnowballOption,
    TernarySnowballOption,
)
from .market import (
    BlackScholesMertonParameters,
    LocalVolatilityParameters,
    LocalVolatilitySurface,
    PricingContext,
    TradingCalendar,
)

from .solvers import (
    ImpliedCouponSettings,
    ImpliedVolatilitySettings,
    SnowballCouponQuote,
    implied_phoenix_coupon,
    implied_snowball_coupon,
    implied_volatility,
)

__version__ = "0.1.0"
