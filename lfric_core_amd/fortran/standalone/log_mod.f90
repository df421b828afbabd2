!> Logging entry points of infrastructure/source/utilities/log_mod.F90:55-60, 386-503 used by the
!> solver API: an ERROR event aborts the run.
module log_mod
  implicit none
  private
  integer, parameter, public :: LOG_LEVEL_ALWAYS = 100000, LOG_LEVEL_ERROR = 200, &
                                LOG_LEVEL_WARNING = 150, LOG_LEVEL_INFO = 100, &
                                LOG_LEVEL_DEBUG = 50, LOG_LEVEL_TRACE = 0
  character(len=512), public :: log_scratch_space
  integer, public :: log_threshold = LOG_LEVEL_INFO
  public :: log_event
contains
  subroutine log_event(message, level)
    character(len=*), intent(in) :: message
    integer, intent(in) :: level
    if (level >= log_threshold) write(*, '(A)') trim(message)
    if (level >= LOG_LEVEL_ERROR .and. level < LOG_LEVEL_ALWAYS) error stop 1
  end subroutine log_event
end module log_mod
